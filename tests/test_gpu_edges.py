"""Edge cases of the ERI plans, Schwarz factors and Fock builds: empty
inputs, bad shell indices, repeated and single-shell quartets, very deep
contractions (one item longer than any split chunk), and a zero density."""
import sys

import numpy as np
import pytest

import libint_amd as la
import oracle
from libint_amd import molecule
from tests.helpers import compare, random_shells

pytestmark = pytest.mark.gpu

ABS, REL = 1e-12, 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_empty_inputs(ctx):
    say("case: empty plan, empty Schwarz list")
    rng = np.random.default_rng(31)
    b = la.Basis(ctx, la.BasisArrays.from_shells(random_shells(rng, [0, 1, 2])))
    plan = la.EriPlan(b, np.zeros((0, 4), np.int32))
    assert plan.size == 0
    plan.execute()
    ctx.synchronize()
    v, nz = plan.download()
    assert v.size == 0 and nz.size == 0
    K = la.schwarz(b, np.zeros((0, 2), np.int32))
    assert np.asarray(K).size == 0


def test_bad_shell_index_is_an_error(ctx):
    say("case: shell index out of range")
    rng = np.random.default_rng(32)
    b = la.Basis(ctx, la.BasisArrays.from_shells(random_shells(rng, [0, 1])))
    for q in ([[0, 1, 0, 2]], [[-1, 0, 0, 0]]):
        with pytest.raises(la.L2gError):
            la.EriPlan(b, np.array(q, np.int32))
    # the context stays usable
    v, _, _ = la.compute_quartets(b, [[0, 1, 0, 1]])
    assert v.size == 9


def test_repeated_and_single_shell_quartets(ctx):
    """(aa|aa) for every l, and the same quartet many times in one batch."""
    rng = np.random.default_rng(33)
    lmax = la.max_l()
    shells = random_shells(rng, list(range(lmax + 1)), kmax=4)
    n = len(shells)
    q = np.array([[i, i, i, i] for i in range(n)] * 3 + [[n - 1, 0, n - 1, 0]] * 200, np.int32)
    say(f"case: single-shell and repeated quartets, nq={len(q)}")
    arrays = la.BasisArrays.from_shells(shells)
    b = la.Basis(ctx, arrays)
    g, goff, gnz = la.compute_quartets(b, q)
    o, ooff, onz = oracle.eri_quartets(arrays, q)
    assert np.array_equal(goff, ooff) and np.array_equal(gnz, onz)
    assert compare(g, o, ABS, REL)[0] == 0
    # repeated quartets give bit-identical blocks
    blk = [g[goff[k]:goff[k] + goff[k + 1] - goff[k]] for k in range(3 * n, 3 * n + 5)]
    assert all(np.array_equal(blk[0], x) for x in blk[1:])


def test_deep_contraction(ctx):
    """K = 16 on every shell of a (dd|dd)-containing set: 65,536 primitive
    quartets in one quartet, Schwarz pairs of 256 primitive pairs (a wave
    per pair), and a Fock build whose items are cut into many
    primitive-range pieces."""
    rng = np.random.default_rng(34)
    shells = []
    # exponents within [0.3, 10]: two d shells of very different widths on
    # different centres put the oracle, which keeps the caller's order, ~1e-10
    # off (HRR cancellation); the GPU builds on the tighter shell
    for l, org in ((2, (0.0, 0.0, 0.0)), (2, (0.0, 0.0, 1.4)), (1, (1.1, 0.3, 0.2))):
        a = np.exp(rng.uniform(np.log(0.3), np.log(10.0), 16))
        shells.append(la.Shell(l, list(zip(a, rng.uniform(0.1, 1.0, 16))), org, pure=True))
    arrays = la.BasisArrays.from_shells(shells)
    b = la.Basis(ctx, arrays)
    q = np.array([[0, 1, 0, 1], [0, 0, 1, 1], [2, 0, 1, 0]], np.int32)
    say("case: K=16 quartets")
    g, _, _ = la.compute_quartets(b, q)
    o, _, _ = oracle.eri_quartets(arrays, q)
    assert compare(g, o, ABS, REL)[0] == 0
    say("case: K=16 Schwarz and Fock")
    pairs = np.array([[0, 0], [1, 0], [1, 1], [2, 2]], np.int32)
    K = la.schwarz(b, pairs)
    _, Ko = oracle.fock_pairs(arrays)
    Ko_p = np.array([Ko[i, j] for i, j in pairs])
    assert compare(np.asarray(K) ** 2, Ko_p ** 2, ABS, REL)[0] == 0
    # and K^2 = max |(ab|ab)| of the GPU's own ERI blocks
    qs = np.array([[i, j, i, j] for i, j in pairs], np.int32)
    gq, goff, _ = la.compute_quartets(b, qs, precision=0.0)
    ends = np.append(goff[1:], gq.size)
    kmax = np.array([np.abs(gq[a:e]).max() for a, e in zip(goff, ends)])
    rel = np.abs(np.asarray(K) ** 2 - kmax) / kmax
    say(f"  K^2 vs max|(ab|ab)|: max rel {rel.max():.2e}")
    # the wave-per-pair Schwarz kernel sums the 65,536 primitive quartets in
    # another order than the ERI kernel does
    assert rel.max() < 1e-11, rel
    N = int(arrays.sizes().sum())
    C = rng.normal(0.0, 0.1, (N, 3))
    D = C @ C.T
    G, J, Kx = la.FockBuilder(b).build(D, precision=1e-12)
    Go, Jo, Ko2, _ = oracle.fock_2e(arrays, D, precision=1e-12)
    assert np.abs(G - Go).max() < 1e-8
    say(f"  |G - G_oracle| {np.abs(G - Go).max():.2e}")


def test_zero_density_fock(ctx):
    say("case: Fock build with D = 0")
    mol = molecule.make_molecule(1, natoms=3)
    b = la.Basis(ctx, mol.basis)
    fb = la.FockBuilder(b)
    D = np.zeros((mol.nbf, mol.nbf))
    G, J, K = fb.build(D, precision=1e-12)
    assert not G.any() and not J.any() and not K.any()
    # and a real density afterwards on the same builder
    rng = np.random.default_rng(35)
    C = rng.normal(0.0, 0.1, (mol.nbf, 4))
    D = C @ C.T
    G, _, _ = fb.build(D, precision=1e-12)
    Go, _, _, _ = oracle.fock_2e(mol.basis, D, precision=1e-12)
    assert np.abs(G - Go).max() < 1e-8


def test_deriv_and_onebody_edges(ctx):
    """Derivative plans: empty, an f shell (raised beyond the kernels' l) and
    a bad index are handled; one-body integrals without charges leave V
    untouched and give symmetric S and T."""
    say("case: derivative plan and one-body edge cases")
    rng = np.random.default_rng(36)
    shells = random_shells(rng, [0, 1, 2, 3])
    arrays = la.BasisArrays.from_shells(shells)
    b = la.Basis(ctx, arrays)
    p = la.Deriv1Plan(b, np.zeros((0, 4), np.int32))
    p.execute()
    ctx.synchronize()
    v, nz = p.download()
    assert v.size == 0 and nz.size == 0
    with pytest.raises(la.L2gError):
        la.Deriv1Plan(b, [[3, 0, 0, 0]])
    with pytest.raises(la.L2gError):
        la.Deriv1Plan(b, [[0, 0, 0, 9]])
    # the context stays usable
    g, _, _ = la.compute_quartets(b, [[2, 1, 2, 0]])
    assert g.size == 5 * 3 * 5
    S, T, V = la.one_body(b, [])
    n = int(arrays.sizes().sum())
    assert S.shape == (n, n) and not V.any()
    assert np.allclose(S, S.T, atol=1e-14) and np.allclose(T, T.T, atol=1e-13)
