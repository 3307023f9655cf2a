"""First geometric derivatives (deriv_order = 1): the oracle's restatement
(oracle/deriv1.py, raised and lowered ERIs combined per Cartesian component)
against Richardson-extrapolated central differences of the oracle's
integrals in the shell centres (an independent route), and for
translational invariance (the 12 sets sum to zero per axis); the GPU run
(test_gpu_eri.py) then checks the device derivative plan against it."""
import numpy as np
import pytest

import libint_amd as la
import oracle
import oracle.deriv1
from tests.helpers import random_shells


def fd_deriv1(arrays, quartet, h=2e-3):
    """d/dX of (s0 s1|s2 s3) by central differences with one Richardson
    step, moving only the k-th shell of the quartet."""
    sh = list(quartet)
    off = np.concatenate([[0], np.cumsum(arrays.nprim)])

    def shell_arrays(shifts):
        ls, pure, np_, al, co, org = [], [], [], [], [], []
        for k, s in enumerate(sh):
            ls.append(int(arrays.l[s]))
            pure.append(int(arrays.pure[s]))
            np_.append(int(arrays.nprim[s]))
            al.append(arrays.alpha[off[s]:off[s + 1]])
            co.append(arrays.coeff[off[s]:off[s + 1]])
            org.append(arrays.origin[3 * s:3 * s + 3] + shifts[k])
        return la.BasisArrays(ls, pure, np_, np.concatenate(al), np.concatenate(co), np.concatenate(org))

    def val(k, d, x):
        sh_ = [np.zeros(3) for _ in range(4)]
        sh_[k][d] = x
        v, _, _ = oracle.eri_quartets(shell_arrays(sh_), [[0, 1, 2, 3]], precision=0.0)
        return v

    out = []
    for k in range(4):
        for d in range(3):
            d1 = (val(k, d, h) - val(k, d, -h)) / (2 * h)
            d2 = (val(k, d, h / 2) - val(k, d, -h / 2)) / h
            out.append((4 * d2 - d1) / 3)
    return np.array(out)


def check_against_fd(arrays, quartets, tol):
    g, offs, nz = oracle.deriv1.eri_deriv1(arrays, quartets)
    sizes = arrays.sizes()
    for i, qq in enumerate(quartets):
        n = int(np.prod(sizes[list(qq)]))
        mine = g[:, offs[i]:offs[i] + n]
        ref = fd_deriv1(arrays, qq)
        assert np.abs(mine - ref).max() < tol, (qq, np.abs(mine - ref).max())
        # translational invariance: the four centres' derivatives cancel per axis
        for d in range(3):
            assert np.abs(mine[d] + mine[3 + d] + mine[6 + d] + mine[9 + d]).max() < 1e-12


def test_deriv1_reference_case():
    """The two contracted Cartesian p shells of the reference's derivative
    test (tests/unit/test-2body.cc:191-196), every quartet over them."""
    p1 = la.Shell(1, [(1.0, 0.9), (0.3, 0.3)], (0.0, 0.0, 0.0), pure=False)
    p2 = la.Shell(1, [(2.0, 0.8), (0.4, -0.2)], (1.0, 1.0, 1.0), pure=False)
    arrays = la.BasisArrays.from_shells([p1, p2])
    q = [(a, b, c, d) for a in range(2) for b in range(2) for c in range(2) for d in range(2)]
    check_against_fd(arrays, q, 1e-8)


def test_deriv1_mixed_pure():
    rng = np.random.default_rng(71)
    shells = random_shells(rng, [0, 1, 2, 2], kmax=2)
    shells[3].pure = False
    arrays = la.BasisArrays.from_shells(shells)
    q = [(2, 1, 3, 0), (0, 0, 1, 1), (3, 2, 2, 1), (1, 3, 0, 2)]
    check_against_fd(arrays, q, 1e-8)
