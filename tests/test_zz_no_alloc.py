"""The no-allocation quartet entry point (l2g_quartet_compute_noalloc)
against the quartet engine it shares its kernels with (l2g_quartet_compute):
bit for bit, on calls in a row whose inputs shrink (the library's scratch
array keeps the previous call's contents), for the three operators and with
precomputed primitive-pair data; sizes beyond its limits fail with an error
that names the limit."""
import sys

import numpy as np
import pytest

import libint_amd as la
from tests.helpers import random_shells

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def shells_of(rng, ls, ks):
    out = []
    for l, k in zip(ls, ks):
        a = rng.uniform(0.1, 8.0, k)
        c = rng.uniform(0.1, 1.0, k)
        out.append(la.Shell(l, list(zip(a, c)), tuple(rng.uniform(0.5, 1.5, 3)), pure=(l >= 2)))
    return out


def check_pair(qe, ctx, s, **kw):
    v, nz = la.quartet_compute_noalloc(ctx, s, **kw)
    r, rnz = qe.compute(s, **kw)
    assert nz == rnz
    assert v.shape == r.shape and np.array_equal(v, r), kw
    return v


def test_noalloc_matches_engine_shrinking_calls(ctx):
    rng = np.random.default_rng(707)
    qe = la.QuartetEngine(ctx, la.NOALLOC_MAX_NPRIM, 3)
    # largest first: (ff|ff) at the primitive limit (generic kernel), then
    # register-kernel classes of fewer primitives and lower l
    cases = [([3, 3, 3, 3], [la.NOALLOC_MAX_NPRIM, 3, 2, 2]),
             ([2, 2, 2, 2], [4, 3, 3, 2]),
             ([1, 0, 2, 1], [2, 1, 1, 2]),
             ([0, 0, 0, 0], [1, 1, 1, 1])]
    for ls, ks in cases:
        s = shells_of(rng, ls, ks)
        for oper, omega in ((la.COULOMB, 0.0), (la.ERF_COULOMB, 0.7), (la.ERFC_COULOMB, 0.7)):
            say(f"case: no-alloc vs engine, l {ls} K {ks} operator {oper}")
            v = check_pair(qe, ctx, s, oper=oper, omega=omega)
            assert np.isfinite(v).all()
    # random quartets of random sizes, each call followed by a smaller one
    for t in range(12):
        big = random_shells(rng, list(rng.integers(0, 4, 4)), kmax=6)
        small = random_shells(rng, list(rng.integers(0, 2, 4)), kmax=2)
        say(f"case: no-alloc pair {t}, l {[x.l for x in big]} then {[x.l for x in small]}")
        check_pair(qe, ctx, big, scale=0.75)
        check_pair(qe, ctx, small, scale=0.75)
    qe._release()


def test_noalloc_given_pairs_and_screening(ctx):
    rng = np.random.default_rng(708)
    qe = la.QuartetEngine(ctx, la.NOALLOC_MAX_NPRIM, 3)
    s = shells_of(rng, [2, 1, 3, 0], [5, 3, 4, 2])
    ln = []
    for a, c in ((s[0], s[1]), (s[2], s[3])):
        AB2 = float(((a.O - c.O) ** 2).sum())
        g = a.alpha[:, None] + c.alpha[None, :]
        lns = -(a.alpha[:, None] * c.alpha[None, :] / g) * AB2 + np.log(np.abs(a.coeff))[:, None] + \
            np.log(np.abs(c.coeff))[None, :]
        ln.append(lns.reshape(-1))
    ln = np.concatenate(ln)
    say("case: no-alloc with precomputed pair data")
    check_pair(qe, ctx, s, precision=1e-12, screening=la.SCREEN_SCHWARZ, ln_scr=ln)
    dropped = ln.copy()
    dropped[: 5 * 3] = -np.inf
    v, nz = la.quartet_compute_noalloc(ctx, s, precision=1e-12, screening=la.SCREEN_SCHWARZ, ln_scr=dropped)
    assert not nz and not v.any()
    check_pair(qe, ctx, s, precision=1e-10, screening=la.SCREEN_CONSERVATIVE)
    qe._release()


def test_noalloc_rejects_sizes_beyond_limits(ctx):
    rng = np.random.default_rng(709)
    too_deep = shells_of(rng, [1, 0, 0, 0], [la.NOALLOC_MAX_NPRIM + 1, 1, 1, 1])
    with pytest.raises(la.L2gError, match="max_nprim 16"):
        la.quartet_compute_noalloc(ctx, too_deep)
    if la.max_l() > 3:
        too_high = shells_of(rng, [4, 0, 0, 0], [1, 1, 1, 1])
        with pytest.raises(la.L2gError, match="max_l 3"):
            la.quartet_compute_noalloc(ctx, too_high)
    # the entry point still works after a rejected call
    s = shells_of(rng, [1, 1, 0, 0], [2, 2, 1, 1])
    qe = la.QuartetEngine(ctx, 2, 1)
    check_pair(qe, ctx, s)
    qe._release()
