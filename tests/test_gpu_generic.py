"""The generic (LDS, table-driven) kernel: every class with an f shell runs
on it; forced on, it must also reproduce the s/p/d classes.  Checked against
the oracle at the integral bar 1e-12 absolute / 1e-10 relative (values the
oracle itself gets wrong are settled by the long-double brute force), and the
Fock builds at 1e-8."""
import itertools
import sys

import numpy as np
import pytest

import libint_amd as la
import oracle
from libint_amd import molecule
from tests.helpers import compare_arbitrated, random_shells

pytestmark = pytest.mark.gpu

ABS, REL = 1e-12, 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_vs_oracle(ctx, arrays, q, **kw):
    b = la.Basis(ctx, arrays)
    g, goff, gnz = la.compute_quartets(b, q, **kw)
    o, ooff, onz = oracle.eri_quartets(arrays, q, **kw)
    assert np.array_equal(goff, ooff)
    return g, o, goff, gnz, onz


def test_generic_forced_on_spd_classes(ctx):
    rng = np.random.default_rng(31)
    shells = random_shells(rng, [0, 1, 2, 0, 1, 2], kmax=3)
    for i, sh in enumerate(shells):
        if sh.l >= 1 and rng.random() < 0.4:
            sh.pure = not sh.pure
    arrays = la.BasisArrays.from_shells(shells)
    q = np.array(list(itertools.product(range(6), repeat=4)), np.int32)
    say(f"case: generic kernel forced on s/p/d classes, {len(q)} quartets")
    ctx.set_generic_kernels(True)
    try:
        g, o, off, gnz, onz = gpu_vs_oracle(ctx, arrays, q)
    finally:
        ctx.set_generic_kernels(False)
    nfail, narb, oerr = compare_arbitrated(arrays, q, off, g, o, ABS, REL)
    say(f"  max abs {np.abs(g - o).max():.2e}; fails {nfail}, arbitrated {narb}")
    assert nfail == 0
    assert np.array_equal(gnz, onz)


@pytest.mark.parametrize("seed", [3, 4])
def test_f_classes_all_orders(ctx, seed):
    """All canonical classes with f shells, in every caller order over a
    set of five shells (s, p, d, f, f), contracted, mixed purity."""
    rng = np.random.default_rng(seed)
    shells = random_shells(rng, [3, 2, 1, 0, 3], kmax=2)
    shells[4].pure = False
    arrays = la.BasisArrays.from_shells(shells)
    q = np.array(list(itertools.product(range(5), repeat=4)), np.int32)
    say(f"case: f classes all orders seed={seed}, {len(q)} quartets")
    g, o, off, gnz, onz = gpu_vs_oracle(ctx, arrays, q)
    nfail, narb, oerr = compare_arbitrated(arrays, q, off, g, o, ABS, REL)
    say(f"  max abs {np.abs(g - o).max():.2e}; fails {nfail}, arbitrated {narb} (oracle error there {oerr:.1e})")
    assert nfail == 0
    assert np.array_equal(gnz, onz)


def test_f_gaussian_order_and_uniform_normalization(ctx):
    rng = np.random.default_rng(8)
    shells = random_shells(rng, [3, 1, 3, 2], kmax=2)
    shells[2].pure = False
    arrays = la.BasisArrays.from_shells(shells)
    q = np.array(list(itertools.product(range(4), repeat=4)), np.int32)[::3]
    say("case: f classes, Gaussian ordering + uniform normalization")
    ctx.set_solid_harmonics_ordering(la.SHG_GAUSSIAN)
    oracle.set_shg_ordering(True)
    try:
        b = la.Basis(ctx, arrays)
        g, goff, _ = la.compute_quartets(b, q, uniform_cart=True)
        o, ooff, _ = oracle.eri_quartets(arrays, q, uniform_cart=True)
        # the arbiter evaluates in the same (Gaussian) ordering
        nfail, _, _ = compare_arbitrated(arrays, q, goff, g, o, ABS, REL, uniform_cart=True)
    finally:
        ctx.set_solid_harmonics_ordering(la.SHG_STANDARD)
        oracle.set_shg_ordering(False)
    assert nfail == 0


def test_f_workload3_sample(ctx):
    """A seeded sample of the unique quartets of the workload-3 recipe
    (K ~ U{1..8}, exponents [0.03, 2e4]) on 2 atoms, all classes."""
    mol = molecule.make_molecule(3, natoms=2)
    B = mol.basis
    allq = molecule.unique_quartets(B.nshell)
    rng = np.random.default_rng(6)
    q = allq[rng.choice(len(allq), min(3000, len(allq)), replace=False)]
    ls = B.l[q]
    say(f"case: workload-3 recipe 2 atoms, {len(q)} quartets, {int((ls == 3).any(axis=1).sum())} with f")
    g, o, off, gnz, onz = gpu_vs_oracle(ctx, B, q)
    nfail, narb, oerr = compare_arbitrated(B, q, off, g, o, ABS, REL)
    say(f"  max abs {np.abs(g - o).max():.2e}; fails {nfail}, arbitrated {narb} (oracle error there {oerr:.1e})")
    assert nfail == 0


def test_f_schwarz_factors(ctx):
    mol = molecule.make_molecule(3, natoms=2)
    B = mol.basis
    b = la.Basis(ctx, B)
    i, j = np.tril_indices(B.nshell)
    pairs = np.stack([i, j], 1)
    say(f"case: Schwarz factors with f shells, {len(pairs)} pairs")
    K = la.schwarz(b, pairs)
    q = np.stack([i, j, i, j], 1).astype(np.int32)
    o, off, _ = oracle.eri_quartets(B, q, precision=0.0)
    sizes = B.sizes()
    n = sizes[i] * sizes[j] * sizes[i] * sizes[j]
    Ko = np.array([np.sqrt(np.abs(o[a:a + m]).max()) for a, m in zip(off, n)])
    assert np.allclose(K ** 2, Ko ** 2, rtol=1e-10, atol=1e-12)
    Kb, S = la.precision_bounds(b, pairs)
    assert np.array_equal(Kb, K)
    assert np.isfinite(S).all() and (S >= 0).all() and (S[K > 1e-100] > 0).all()


@pytest.mark.parametrize("generic_all", [False, True])
def test_f_fock_vs_oracle(ctx, generic_all):
    mol = molecule.make_molecule(3, natoms=2)
    rng = np.random.default_rng(14)
    C = rng.normal(0.0, 0.1, (mol.nbf, max(1, mol.nbf // 5)))
    D = C @ C.T
    say(f"case: Fock with f shells nbf={mol.nbf}, generic kernel for all classes: {generic_all}")
    ctx.set_generic_kernels(generic_all)
    try:
        b = la.Basis(ctx, mol.basis)
        fb = la.FockBuilder(b)
        G, J, K = fb.build(D, 1e-12)
    finally:
        ctx.set_generic_kernels(False)
    Go, Jo, Ko, _ = oracle.fock_2e(mol.basis, D, 1e-12)
    say(f"  max|G-Go| {np.abs(G - Go).max():.2e} |J-Jo| {np.abs(J - Jo).max():.2e} |K-Ko| {np.abs(K - Ko).max():.2e}")
    assert np.abs(G - Go).max() <= 1e-8
    assert np.abs(J - Jo).max() <= 1e-8
    assert np.abs(K - Ko).max() <= 1e-8


def test_few_deep_quartets_route_to_generic(ctx):
    """With the wave split off, a class with a register kernel, total
    L >= 7, few quartets and deep contractions runs on the generic kernel (a
    block per quartet, each contraction in its serial order), at the
    integral bar; with the split on (the default) it runs a wave per
    quartet, or the generic kernel where that was measured faster for the
    class (mode 2 keeps it on the wave); the same class with shallow
    contractions stays on its register kernel."""
    rng = np.random.default_rng(41)
    deep = []
    for k in range(4):
        a = np.exp(rng.uniform(np.log(0.3), np.log(20.0), 6))
        deep.append(la.Shell(2, list(zip(a, rng.uniform(0.2, 1.0, 6))), tuple(rng.normal(0, 0.8, 3)), pure=True))
    arrays = la.BasisArrays.from_shells(deep)
    q = np.array(list(itertools.product(range(4), repeat=4)), np.int32)[:60]
    say(f"case: {len(q)} (dd|dd) quartets with K = 6 on every shell")
    b = la.Basis(ctx, arrays)
    plan = la.EriPlan(b, q)
    st = plan.stats()
    assert [(s["cls"], s["generic"] or s["wave"]) for s in st] == [("2222", True)], st
    plan.set_wave_split(2)
    st = plan.stats()
    assert [(s["cls"], s["generic"], s["wave"]) for s in st] == [("2222", False, True)], st
    plan.set_wave_split(False)
    st = plan.stats()
    assert [(s["cls"], s["generic"]) for s in st] == [("2222", True)], st
    plan.execute()
    g, gnz = plan.download()
    o, off, onz = oracle.eri_quartets(arrays, q)
    nfail, narb, oerr = compare_arbitrated(arrays, q, off, g, o, ABS, REL)
    say(f"  max abs {np.abs(g - o).max():.2e}; fails {nfail}, arbitrated {narb}")
    assert nfail == 0
    assert np.array_equal(gnz, onz)
    shallow = random_shells(rng, [2, 2, 2, 2], kmax=1)
    st2 = la.EriPlan(la.Basis(ctx, la.BasisArrays.from_shells(shallow)), q).stats()
    assert [(s["cls"], s["generic"]) for s in st2] == [("2222", False)], st2


def test_deep_quartets_split_over_blocks(ctx):
    """Few (fd|ff) and (ff|ff) quartets of 8-primitive shells (4096 primitive
    quartets each): the generic kernel splits each quartet's primitive loop
    into pieces run by separate blocks and sums them per quartet in piece
    order.  At the integral bar against the oracle, and the same bits on a
    second execution (the sum has a fixed order)."""
    rng = np.random.default_rng(77)
    shells = [la.Shell(l, list(zip(np.exp(rng.uniform(np.log(0.05), np.log(5e3), 8)), rng.uniform(0.1, 1.0, 8))),
                       tuple(rng.uniform(-0.6, 0.6, 3)), pure=p)
              for l, p in ((3, True), (2, True), (3, False), (3, True))]
    arrays = la.BasisArrays.from_shells(shells)
    q = np.array([(0, 1, 2, 3), (2, 3, 0, 1), (0, 2, 3, 0), (3, 2, 2, 0), (0, 3, 2, 3), (1, 0, 3, 2),
                  (0, 0, 3, 3), (2, 0, 3, 1)], np.int32)
    say(f"case: {len(q)} deep f quartets (K = 8 on every shell), split primitive loops")
    b = la.Basis(ctx, arrays)
    plan = la.EriPlan(b, q)
    plan.execute()
    g1, nz1 = plan.download()
    plan.execute()
    g2, _ = plan.download()
    assert np.array_equal(g1, g2)
    o, off, onz = oracle.eri_quartets(arrays, q)
    assert np.array_equal(plan.offsets(), off)
    g1 = g1[:len(o)]
    nfail, narb, oerr = compare_arbitrated(arrays, q, off, g1, o, ABS, REL)
    say(f"  classes {[c['cls'] + ('/generic' if c['generic'] else '') for c in plan.stats()]}; "
        f"max abs {np.abs(g1 - o).max():.2e}; fails {nfail}, arbitrated {narb}")
    assert nfail == 0
