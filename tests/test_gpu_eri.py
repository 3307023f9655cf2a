"""GPU parity of the ERI batch path against the CPU oracle.

Bar: per integral, fail only if |gpu - oracle| > 1e-12 AND the relative error
> 1e-10 (the reference's own rule shape is
tests/eri/test.cc:77-88).  Every call goes through the C ABI (libl2g.so).
"""
import itertools
import sys

import numpy as np
import pytest

import libint_amd as la
import oracle
import oracle.deriv1
from libint_amd import molecule
from tests.helpers import compare_arbitrated, compare, random_shells, within_reference

pytestmark = pytest.mark.gpu

ABS, REL = 1e-12, 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_vs_oracle(ctx, shells_or_arrays, quartets, **kw):
    arrays = shells_or_arrays if isinstance(shells_or_arrays, la.BasisArrays) else \
        la.BasisArrays.from_shells(shells_or_arrays)
    b = la.Basis(ctx, arrays)
    prec = kw.get("precision", np.finfo(float).eps)
    scr = kw.get("screening", la.SCREEN_ORIGINAL)
    scale = kw.get("scale", 1.0)
    uni = kw.get("uniform_cart", False)
    g, goff, gnz = la.compute_quartets(b, quartets, prec, scr, scale, uni)
    o, ooff, onz = oracle.eri_quartets(arrays, quartets, precision=prec, screening=scr, uniform_cart=uni)
    if scale != 1.0:
        o = o * scale
    assert np.array_equal(goff, ooff)
    return g, o, gnz, onz


def test_python_goldens_gpu(ctx):
    say("case: python goldens (pp|ss), [s,p,s,p]")
    s = la.Shell(0, [(1.0, 10.0)])
    p = la.Shell(1, [(1.0, 10.0)])
    b = la.Basis(ctx, la.BasisArrays.from_shells([s, p]))
    v, _, _ = la.compute_quartets(b, [[1, 1, 0, 0]])
    assert abs(np.linalg.norm(v) - 1.62867503968) < 1e-10
    b4 = la.Basis(ctx, la.BasisArrays.from_shells([s, p, s, p]))
    q = list(itertools.product(range(4), repeat=4))
    v, _, _ = la.compute_quartets(b4, q)
    assert abs(np.linalg.norm(v) - 14.7036075402) < 1e-9


@pytest.mark.parametrize("seed", [1, 2])
def test_all_classes_all_orders(ctx, seed):
    """Every class up to max_l, every caller order of the four shells,
    mixed purity, contracted (K<=3)."""
    lmax = la.max_l()
    rng = np.random.default_rng(seed)
    ls = [l for l in range(lmax + 1) for _ in range(2)]
    shells = random_shells(rng, ls, kmax=3)
    # flip purity of some p/d shells
    for i, sh in enumerate(shells):
        if sh.l >= 1 and rng.random() < 0.4:
            sh.pure = not sh.pure
    n = len(shells)
    q = np.array(list(itertools.product(range(n), repeat=4)), np.int32)
    say(f"case: all classes all orders seed={seed} nshell={n} nq={len(q)}")
    g, o, gnz, onz = gpu_vs_oracle(ctx, shells, q)
    nfail, mabs, mrel = compare(g, o, ABS, REL)
    say(f"  max abs {mabs:.3e} max rel {mrel:.3e} fails {nfail}")
    assert nfail == 0
    assert np.array_equal(gnz, onz)


def test_config1_dddd_and_sample(ctx):
    """Workload 1 (10 atoms, 6 A cube): every (dd|dd) quartet plus a seeded
    sample of the other unique quartets, against the oracle."""
    mol = molecule.make_molecule(1)
    B = mol.basis
    allq = molecule.unique_quartets(B.nshell)
    ls = B.l[allq]
    dd = allq[(ls == 2).all(axis=1)]
    rng = np.random.default_rng(7)
    other = allq[rng.choice(len(allq), 40000, replace=False)]
    q = np.concatenate([dd, other])
    say(f"case: workload-1 (dd|dd)={len(dd)} + sample={len(other)}")
    g, o, gnz, onz = gpu_vs_oracle(ctx, B, q)
    nfail, mabs, mrel = compare(g, o, ABS, REL)
    say(f"  max abs {mabs:.3e} max rel {mrel:.3e} fails {nfail}")
    assert nfail == 0
    assert np.array_equal(gnz, onz)


def test_gaussian_order_uniform_prescale(ctx):
    rng = np.random.default_rng(11)
    shells = random_shells(rng, [2, 1, 2, 0, 1], kmax=2)
    q = np.array(list(itertools.product(range(5), repeat=4)), np.int32)[::3]
    say("case: gaussian ordering + uniform normalization + prescale")
    ctx.set_solid_harmonics_ordering(la.SHG_GAUSSIAN)
    oracle.set_shg_ordering(True)
    try:
        g, o, _, _ = gpu_vs_oracle(ctx, shells, q, uniform_cart=True, scale=1.7)
    finally:
        ctx.set_solid_harmonics_ordering(la.SHG_STANDARD)
        oracle.set_shg_ordering(False)
    nfail, mabs, mrel = compare(g, o, ABS, REL)
    assert nfail == 0


def test_screening_precision_and_null_results(ctx):
    """Distant tight shells at a loose precision: quartets whose every
    primitive is screened return nullptr in libint (engine.impl.h:1781-1784);
    the GPU flags them identically and writes zeros."""
    s1 = la.Shell(0, [(50.0, 1.0)], (0, 0, 0))
    s2 = la.Shell(1, [(40.0, 1.0), (3.0, 0.5)], (9.0, 0, 0))
    s3 = la.Shell(2, [(30.0, 1.0)], (0, 9.0, 0))
    s4 = la.Shell(0, [(0.2, 1.0)], (0, 0, 1.0))
    shells = [s1, s2, s3, s4]
    q = np.array(list(itertools.product(range(4), repeat=4)), np.int32)
    for prec, scr in [(1e-10, la.SCREEN_ORIGINAL), (1e-8, la.SCREEN_CONSERVATIVE), (0.0, la.SCREEN_ORIGINAL)]:
        say(f"case: screening prec={prec} method={scr}")
        g, o, gnz, onz = gpu_vs_oracle(ctx, shells, q, precision=prec, screening=scr)
        nfail, _, _ = compare(g, o, ABS, REL)
        assert nfail == 0
        assert np.array_equal(gnz, onz)
        if prec > 0:
            assert (gnz == 0).any()


def test_engine_facade_single_quartet(ctx):
    rng = np.random.default_rng(5)
    sh = random_shells(rng, [2, 1, 2, 2], kmax=3)
    say("case: Engine.compute single quartet")
    e = la.Engine(max_nprim=3, max_l=2, ctx=ctx)
    v = e.compute(*sh)
    o, _, _ = oracle.eri_quartets(la.BasisArrays.from_shells(sh), [[0, 1, 2, 3]])
    nfail, _, _ = compare(v, o, ABS, REL)
    assert nfail == 0


def test_schwarz_factors(ctx):
    """GPU Schwarz kernels vs max|.| / ||.||_2 of the oracle's (ab|ab) at
    precision 0 (tests/hartree-fock/hartree-fock++.cc:1231-1298)."""
    mol = molecule.make_molecule(1)
    B = mol.basis
    i, j = np.tril_indices(B.nshell)
    pairs = np.stack([i, j], 1).astype(np.int32)
    say(f"case: Schwarz factors workload-1, {len(pairs)} pairs")
    b = la.Basis(ctx, B)
    Kinf = la.schwarz(b, pairs, 0.0, "inf")
    K2 = la.schwarz(b, pairs, 0.0, "2")
    q = np.stack([pairs[:, 0], pairs[:, 1], pairs[:, 0], pairs[:, 1]], 1)
    o, off, _ = oracle.eri_quartets(B, q, precision=0.0)
    ends = np.append(off[1:], o.size)
    Oinf = np.array([np.sqrt(np.abs(o[a:e]).max()) for a, e in zip(off, ends)])
    O2 = np.array([np.sqrt(np.linalg.norm(o[a:e])) for a, e in zip(off, ends)])
    # K is the square root of an integral norm; parity is held on the norm
    # itself (an integral-valued quantity) at the per-integral bar, since the
    # square root amplifies the ~1e-17 absolute noise of tiny norms
    assert compare(Kinf ** 2, Oinf ** 2, ABS, REL)[0] == 0
    assert compare(K2 ** 2, O2 ** 2, ABS, REL)[0] == 0


def test_context_close_releases_children():
    """Closing a context with a basis and a plan still referenced releases
    them first; the C ABI itself refuses to finalize a context with live
    objects (L2G_E_STATE) instead of leaving them dangling."""
    say("case: context lifetime")
    c = la.Context(0)
    rng = np.random.default_rng(3)
    b = la.Basis(c, la.BasisArrays.from_shells(random_shells(rng, [0, 1])))
    plan = la.EriPlan(b, [[0, 1, 0, 1]])
    plan.execute()
    rc = la.lib().l2g_finalize(c.handle)
    assert rc == la.L2G_E_STATE
    c.close()
    with pytest.raises(la.L2gError):
        plan.execute()
    with pytest.raises(la.L2gError):
        b.handle


def test_graph_replay_matches_eager(ctx):
    """Graph mode: the captured class launches (specialized FP64 and FP32
    kernels, and the generic kernel for f shells) replayed several times,
    with eager executions of other plans overwriting the output between the
    replays, give the eager result bit for bit every time; a new output
    pointer re-captures."""
    rng = np.random.default_rng(21)
    lmax = la.max_l()
    shells = random_shells(rng, [l for l in range(lmax + 1) for _ in range(2)], kmax=2)
    n = len(shells)
    allq = np.array(list(itertools.product(range(n), repeat=4)), np.int32)
    q = allq[rng.choice(len(allq), min(len(allq), 3000), replace=False)]
    say(f"case: graph replay nshell={n} nq={len(q)}")
    b = la.Basis(ctx, la.BasisArrays.from_shells(shells))
    for thr in (0.0, 1e-4):
        P = la.EriPlan(b, q, fp32_threshold=thr)
        P.execute()
        ref, refnz = P.download()
        kinds = {(s["fp32"], s["generic"]) for s in P.stats()}
        say(f"  fp32_threshold={thr}: kernel kinds {sorted(kinds)}")
        if lmax >= 3:
            assert (False, True) in kinds
        # same quartets reversed: same output size, different layout
        Z = la.EriPlan(b, q[::-1].copy(), fp32_threshold=thr)
        other = la.EriPlan(b, allq[:500], fp32_threshold=thr)
        P.set_graph(True)
        for rep in range(4):
            P.execute()
            ctx.synchronize()
            v, nz = P.download()
            assert np.array_equal(v, ref), f"replay {rep}"
            assert np.array_equal(nz, refnz)
            assert all(s["ms"] == -1 for s in P.stats())
            assert P.elapsed_ms() > 0
            # eager work between replays, overwriting P's output
            Z.execute(P.device_output())
            other.execute()
            ctx.synchronize()
            assert not np.array_equal(P.download()[0], ref)
        # a second output pointer re-captures; then back to the first
        H = la.EriPlan(b, q, fp32_threshold=thr)
        H.execute()
        Z.execute(H.device_output())
        ctx.synchronize()
        assert not np.array_equal(H.download()[0], ref)
        P.execute(H.device_output())
        ctx.synchronize()
        assert np.array_equal(H.download()[0], ref)
        P.execute()
        ctx.synchronize()
        assert np.array_equal(P.download()[0], ref)
        # switching the solid-harmonic ordering re-captures: the replay gives
        # the eager Gaussian-order result, and switching back the Standard one
        E = la.EriPlan(b, q, fp32_threshold=thr)
        ctx.set_solid_harmonics_ordering(la.SHG_GAUSSIAN)
        try:
            E.execute()
            gref, _ = E.download()
            P.execute()
            ctx.synchronize()
            assert np.array_equal(P.download()[0], gref)
        finally:
            ctx.set_solid_harmonics_ordering(la.SHG_STANDARD)
        if any(s.pure and s.l >= 2 for s in shells):
            assert not np.array_equal(gref, ref)
        P.execute()
        ctx.synchronize()
        assert np.array_equal(P.download()[0], ref)
        E._release()
        # graph off: eager again, with per-class timings
        P.set_graph(False)
        Z.execute(P.device_output())
        P.execute()
        ctx.synchronize()
        assert np.array_equal(P.download()[0], ref)
        assert all(s["ms"] >= 0 for s in P.stats())
        for x in (P, Z, other, H):
            x._release()


def test_wave_split(ctx):
    """Wave per quartet for classes of few deep quartets (the default):
    classes left in lane mode are bit-identical to a plan with the wave
    split off; wave classes sum the contraction in another order and meet
    the per-integral bar with quad-precision arbitration, as the
    lane-per-quartet order does."""
    mol = molecule.make_molecule(3, natoms=3, same_density=True)
    B = mol.basis
    b = la.Basis(ctx, B)
    allq = molecule.unique_quartets(B.nshell)
    rng = np.random.default_rng(4)
    q = allq[rng.choice(len(allq), 12000, replace=False)]
    say(f"case: wave split, {len(q)} quartets of the workload-3 recipe (3 atoms)")
    ref = la.EriPlan(b, q)
    ref.set_wave_split(False)
    ref.execute()
    r, rnz = ref.download()
    plan = la.EriPlan(b, q)
    plan.execute()
    g, gnz = plan.download()
    st = plan.stats()
    wave = {s["cls"] for s in st if s["wave"]}
    say(f"  wave classes: {sorted(wave)}")
    assert wave
    assert np.array_equal(gnz, rnz)
    offs = plan.offsets()
    sizes = B.sizes()
    n = sizes[q[:, 0]] * sizes[q[:, 1]] * sizes[q[:, 2]] * sizes[q[:, 3]]
    # canonical class of each quartet: bra pair ordered (la >= lb), the
    # lighter pair first
    ls = B.l[q]
    bra = np.sort(ls[:, :2], axis=1)[:, ::-1]
    ket = np.sort(ls[:, 2:], axis=1)[:, ::-1]
    swap = bra.sum(1) > ket.sum(1)
    c1 = np.where(swap[:, None], ket, bra)
    c2 = np.where(swap[:, None], bra, ket)
    cls = np.array([f"{a}{b_}{c}{d}" for (a, b_), (c, d) in zip(c1, c2)])
    inw = np.isin(cls, list(wave))
    iw = np.repeat(inw, n)
    assert np.array_equal(g[~iw], r[~iw])
    o, ooff, _ = oracle.eri_quartets(B, q)
    det = {}
    nfail, narb, _ = compare_arbitrated(B, q[inw], np.concatenate([[0], np.cumsum(n[inw])[:-1]]), g[iw], o[iw],
                                        1e-12, 1e-10, detail=det)
    say(f"  wave part ({int(inw.sum())} quartets): max abs {np.abs(g[iw] - o[iw]).max():.2e}, fails {nfail}, "
        f"{narb} values arbitrated")
    assert nfail == 0
    assert within_reference(det)


def test_three_and_two_centre(ctx):
    """3-centre (a|cd) and 2-centre (a|c) Coulomb integrals: the 4-centre
    kernels with Shell.unit() (exponent 0, coefficient 1) as the missing
    centre, as Engine::compute does for BraKet xs_xx / xs_xs
    (engine.impl.h:140-178); every class to l = 3 against the oracle, and
    the reference's python golden ||(s s|s)|| = 3.6563211198
    (python/tests/test_libint2.py:39-42)."""
    rng = np.random.default_rng(41)
    lmax = la.max_l()
    shells = random_shells(rng, [l for l in range(lmax + 1) for _ in range(2)], kmax=3)
    n = len(shells)
    u = n
    arrays = la.BasisArrays.from_shells(shells + [la.Shell.unit()])
    q3 = [(a, u, c, d) for a in range(n) for c in range(n) for d in range(n)]
    q2 = [(a, u, c, u) for a in range(n) for c in range(n)]
    q = np.array(q3 + q2, np.int32)
    say(f"case: 3-centre {len(q3)} + 2-centre {len(q2)} sets")
    g, o, gnz, onz = gpu_vs_oracle(ctx, arrays, q)
    nfail, mabs, mrel = compare(g, o, ABS, REL)
    say(f"  max abs {mabs:.3e} max rel {mrel:.3e} fails {nfail}")
    assert nfail == 0
    assert np.array_equal(gnz, onz)
    s = la.Shell(0, [(1.0, 10.0)])
    for bk in ("XSXX", "XXXS"):
        v = la.Engine(1, 0, ctx=ctx, braket=bk).compute(s, s, s)
        assert abs(np.linalg.norm(v) - 3.6563211198) < 1e-9
    v = la.Engine(1, 0, ctx=ctx, braket="XSXS").compute(s, s)
    o2, _, _ = oracle.eri_quartets(la.BasisArrays.from_shells([s, la.Shell.unit()]), [[0, 1, 0, 1]])
    assert compare(v, o2, ABS, REL)[0] == 0


@pytest.mark.parametrize("oper", [la.ERF_COULOMB, la.ERFC_COULOMB])
def test_range_separated(ctx, oper):
    """erf_coulomb and erfc_coulomb (engine.h:158-167) at two attenuations,
    every class to l = 3, against the oracle at the per-integral bar (the
    oracle's core integrals are pinned by quadrature and by the
    McMurchie-Davidson brute force); omega = 0 gives zeros for erf and the
    Coulomb integrals for erfc."""
    rng = np.random.default_rng(61)
    lmax = la.max_l()
    shells = random_shells(rng, [l for l in range(lmax + 1) for _ in range(2)], kmax=2)
    n = len(shells)
    allq = np.array(list(itertools.product(range(n), repeat=4)), np.int32)
    q = allq[rng.choice(len(allq), 1500, replace=False)]
    arrays = la.BasisArrays.from_shells(shells)
    b = la.Basis(ctx, arrays)
    for w in (0.4, 2.5):
        say(f"case: operator {oper} omega {w}, {len(q)} quartets")
        plan = la.EriPlan(b, q)
        plan.set_operator(oper, w)
        plan.execute()
        g, gnz = plan.download()
        o, _, onz = oracle.eri_quartets(arrays, q, oper=oper, omega=w)
        nfail, mabs, mrel = compare(g, o, ABS, REL)
        say(f"  max abs {mabs:.3e} max rel {mrel:.3e} fails {nfail}")
        assert nfail == 0
        assert np.array_equal(gnz, onz)
    plan = la.EriPlan(b, q)
    plan.set_operator(oper, 0.0)
    plan.execute()
    g, _ = plan.download()
    if oper == la.ERF_COULOMB:
        assert not g.any()
    else:
        c, _, _ = la.compute_quartets(b, q)
        assert np.array_equal(g, c)


def test_first_derivatives(ctx):
    """deriv_order = 1 (12 shell sets per quartet, libint's results()[0..11])
    from the device derivative plan: its raised and lowered ERIs and its
    combined shell sets against the oracle's (oracle/deriv1.py, itself
    pinned by finite differences in tests/test_deriv.py) at the integral bar,
    the Gaussian solid-harmonics ordering as a permutation of the Standard
    one, translational invariance, and the Engine facade on the reference's
    derivative-test shells (tests/unit/test-2body.cc:191-196)."""
    from libint_amd import deriv
    rng = np.random.default_rng(72)
    shells = random_shells(rng, [0, 1, 2, 1, 2], kmax=3)
    shells[4].pure = False
    arrays = la.BasisArrays.from_shells(shells)
    n = len(shells)
    allq = np.array(list(itertools.product(range(n), repeat=4)), np.int32)
    q = allq[rng.choice(len(allq), 300, replace=False)]
    say(f"case: first derivatives, {len(q)} quartets")
    g, offs, nz = deriv.eri_deriv1(ctx, arrays, q)
    # the raised and lowered ERIs against the oracle's at the integral bar
    aug = oracle.deriv1.augmented(arrays)
    n_ = arrays.nshell
    big = np.concatenate([q + np.eye(4, dtype=np.int32)[k] * n_ for k in range(4)] +
                         [q + np.eye(4, dtype=np.int32)[k] * 2 * n_ for k in range(4)])
    gv, _, _ = la.compute_quartets(la.Basis(ctx, la.BasisArrays(aug.l, aug.pure, aug.nprim, aug.alpha, aug.coeff,
                                                                aug.origin)), big)
    ov, _, _ = oracle.eri_quartets(aug, big)
    assert compare(gv, ov, ABS, REL)[0] == 0
    # the device combination against the oracle's restatement, same bar
    o, ooff, onz = oracle.deriv1.eri_deriv1(arrays, q)
    assert np.array_equal(offs, ooff) and np.array_equal(nz, onz)
    for d in range(12):
        assert compare(g[d], o[d], ABS, REL)[0] == 0, d
    # Gaussian ordering (0, +1, -1, +2, -2) of the pure centres
    ctx.set_solid_harmonics_ordering(la.SHG_GAUSSIAN)
    try:
        gg, _, _ = deriv.eri_deriv1(ctx, arrays, q[:40])
    finally:
        ctx.set_solid_harmonics_ordering(la.SHG_STANDARD)
    sizes = arrays.sizes()
    for i in range(40):
        dims = [int(sizes[s_]) for s_ in q[i]]
        blk = g[:, offs[i]:offs[i] + int(np.prod(dims))].reshape([12] + dims)
        for k in range(4):
            sh = q[i, k]
            if arrays.pure[sh] and arrays.l[sh] > 0:
                l_ = int(arrays.l[sh])
                std = [l_ + ((p_ + 1) // 2 if p_ % 2 else -(p_ // 2)) for p_ in range(2 * l_ + 1)]
                blk = np.take(blk, std, axis=k + 1)
        # another batch (40 quartets, not 300): classes may switch between a
        # lane and a wave per quartet, which changes the rounding only
        x = gg[:, offs[i]:offs[i] + int(np.prod(dims))]
        assert np.abs(x - blk.reshape(12, -1)).max() <= 1e-14 * max(1.0, np.abs(blk).max()), i
    for d in range(3):
        s_ = g[d] + g[3 + d] + g[6 + d] + g[9 + d]
        assert np.abs(s_).max() < 1e-11
    p1 = la.Shell(1, [(1.0, 0.9), (0.3, 0.3)], (0.0, 0.0, 0.0), pure=False)
    p2 = la.Shell(1, [(2.0, 0.8), (0.4, -0.2)], (1.0, 1.0, 1.0), pure=False)
    e = la.Engine(2, 1, ctx=ctx, deriv_order=1)
    sets = e.compute(p1, p2, p1, p2)
    assert len(sets) == 12 and all(s.size == 81 for s in sets)
    for d in range(3):
        assert np.abs(sets[d] + sets[3 + d] + sets[6 + d] + sets[9 + d]).max() < 1e-12
