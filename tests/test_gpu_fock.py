"""GPU screened Fock build (J, K, G = 2J - K) against the oracle's
restatement of compute_2body_fock (tests/hartree-fock/hartree-fock++.cc:
1574-1772).  Fock-matrix error <= 1e-8."""
import sys

import numpy as np
import pytest

import libint_amd as la
import oracle
from libint_amd import molecule

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def density(nbf, seed):
    rng = np.random.default_rng(seed)
    C = rng.normal(0.0, 0.1, (nbf, max(1, nbf // 5)))
    return C @ C.T


@pytest.mark.parametrize("workload,natoms,screening", [
    (1, 5, la.SCREEN_SCHWARZINF),
    (2, 3, la.SCREEN_SCHWARZINF),
    (1, 4, la.SCREEN_ORIGINAL),
])
def test_fock_vs_oracle(ctx, workload, natoms, screening):
    mol = molecule.make_molecule(workload, natoms=natoms)
    D = density(mol.nbf, 10 + natoms)
    say(f"case: fock workload={workload} natoms={natoms} nbf={mol.nbf} screening={screening}")
    b = la.Basis(ctx, mol.basis)
    fb = la.FockBuilder(b, screening=screening)
    G, J, K = fb.build(D, precision=1e-12)
    Go, Jo, Ko, nqo = oracle.fock_2e(mol.basis, D, precision=1e-12, screening=screening)
    st = fb.stats()
    say(f"  quartets gpu {st['nquartets']} oracle {nqo}; max|G-Go| {np.abs(G - Go).max():.2e} "
        f"max|J-Jo| {np.abs(J - Jo).max():.2e} max|K-Ko| {np.abs(K - Ko).max():.2e}")
    assert np.abs(G - Go).max() <= 1e-8
    assert np.abs(J - Jo).max() <= 1e-8
    assert np.abs(K - Ko).max() <= 1e-8
    assert np.allclose(G, G.T, atol=1e-14)


def test_fock_pairs_and_schwarz(ctx):
    mol = molecule.make_molecule(1, natoms=6)
    b = la.Basis(ctx, mol.basis)
    fb = la.FockBuilder(b)
    K = fb.schwarz()
    mask, Ko = oracle.fock_pairs(mol.basis)
    sig = np.tril(mask.astype(bool))
    say(f"case: fock pair list, {sig.sum()} significant pairs of {mask.shape[0] * (mask.shape[0] + 1) // 2}")
    assert np.array_equal(np.tril(K) > 0, sig)
    d = np.abs(K ** 2 - Ko ** 2)[sig]
    assert d.max() <= 1e-12


def test_fock_dense_identity_small(ctx):
    """G = 2J - K against a dense unscreened contraction of the GPU's own
    ERI tensor (hartree-fock.cc:536-626 semantics)."""
    import itertools
    mol = molecule.make_molecule(1, natoms=2)
    B = mol.basis
    N = mol.nbf
    D = density(N, 3)
    b = la.Basis(ctx, B)
    ns = B.nshell
    q = np.array(list(itertools.product(range(ns), repeat=4)), np.int32)
    say(f"case: dense identity nbf={N} quartets={len(q)}")
    v, offs, _ = la.compute_quartets(b, q, precision=0.0)
    sizes = B.sizes()
    off = np.concatenate([[0], np.cumsum(sizes)])
    T = np.zeros((N, N, N, N))
    for k, (a, bb, c, d) in enumerate(q):
        n = sizes[a] * sizes[bb] * sizes[c] * sizes[d]
        T[off[a]:off[a + 1], off[bb]:off[bb + 1], off[c]:off[c + 1], off[d]:off[d + 1]] = \
            v[offs[k]:offs[k] + n].reshape(sizes[a], sizes[bb], sizes[c], sizes[d])
    Jd = np.einsum("abcd,cd->ab", T, D)
    Kd = np.einsum("acbd,cd->ab", T, D)
    G, J, K = la.FockBuilder(b).build(D, precision=1e-13)
    assert np.abs(J - Jd).max() < 1e-9
    assert np.abs(K - Kd).max() < 1e-9
    assert np.abs(G - (2 * Jd - Kd)).max() < 1e-9


def test_fock_deterministic_digest(ctx):
    """Fixed-point J/K accumulation (l2g_fock_set_deterministic) on an
    f-recipe molecule whose build runs register, split deep and generic
    (f) classes: two builds give the same bits; the result meets the oracle
    at the Fock bar and the FP64-atomic build to rounding; the count pass
    alone gives the build's quartet count."""
    mol = molecule.make_molecule(4, natoms=3, same_density=True)
    D = density(mol.nbf, 31)
    say(f"case: deterministic Fock digest, f recipe 3 atoms, nbf {mol.nbf}")
    b = la.Basis(ctx, mol.basis)
    fb = la.FockBuilder(b)
    Gf, Jf, Kf = fb.build(D, 1e-12)
    nq = fb.stats()["nquartets"]
    assert fb.count(D, 1e-12) == nq
    fb.set_deterministic(True)
    runs = [fb.build(D, 1e-12) for _ in range(3)]
    cls = fb.class_stats()
    say(f"  {nq} quartets, classes {sorted({c['cls'] for c in cls})}")
    for r in runs[1:]:
        for x, y in zip(runs[0], r):
            assert np.array_equal(x, y)
    G, J, K = runs[0]
    for x, y in ((G, Gf), (J, Jf), (K, Kf)):
        assert np.abs(x - y).max() <= 1e-13 * max(1.0, np.abs(y).max())
    Go, Jo, Ko, _ = oracle.fock_2e(mol.basis, D, 1e-12, tight_centre=True)
    say(f"  deterministic vs oracle: max|G-Go| {np.abs(G - Go).max():.2e}")
    assert np.abs(G - Go).max() <= 1e-8 and np.abs(J - Jo).max() <= 1e-8 and np.abs(K - Ko).max() <= 1e-8
    # windows of a few hundred quartets: the same sum of the same contributions
    fb.set_window(500)
    Gw, _, _ = fb.build(D, 1e-12)
    assert fb.stats()["windows"] > 1
    assert np.abs(Gw - G).max() <= 1e-13 * max(1.0, np.abs(G).max())
