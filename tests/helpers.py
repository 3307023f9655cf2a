"""Shared helpers for the tests: seeded random shells and the pass rule."""
import numpy as np

import libint_amd as la


def random_shells(rng, ls, kmax=3, box=(0.7, 1.3), expo=(0.1, 3.0), pure=None):
    """Random contracted shells in the spirit of the reference's
    RandomShellSet (src/bin/test_eri/prep_libint2.h:35-69), but seeded."""
    out = []
    for l in ls:
        K = int(rng.integers(1, kmax + 1))
        a = rng.uniform(*expo, K)
        c = rng.uniform(0.1, 1.0, K)
        p = (l >= 2) if pure is None else pure
        out.append(la.Shell(l, list(zip(a, c)), tuple(rng.uniform(*box, 3)), pure=p))
    return out


def compare(got, ref, abs_tol, rel_tol):
    """The reference's rule: a value fails only if BOTH the absolute and the
    relative error exceed their thresholds (tests/eri/test.cc:77-88, 436).
    Returns (nfail, max_abs, max_rel)."""
    got = np.asarray(got, float)
    ref = np.asarray(ref, float)
    d = np.abs(got - ref)
    rel = d / np.maximum(np.abs(ref), 1e-300)
    fail = (d > abs_tol) & (rel > rel_tol)
    return int(fail.sum()), float(d.max(initial=0.0)), float(np.where(np.abs(ref) > abs_tol, rel, 0).max(initial=0.0))
