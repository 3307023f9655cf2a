"""FP32/FP64 mixed precision: quartets with Schwarz bound K12 K34 < 1e-6 and
pre-HRR bound S12 S34 < 16e-6 run in FP32, the rest in FP64.  Errors are
reported per class against the FP64 oracle; FP64 quartets keep the 1e-12 /
1e-10 bar, FP32 quartets must stay within 1e-11 absolute (1e-5 of the
threshold) and within 64 eps_fp32 S12 S34, and the mixed Fock matrix within
1e-8."""
import sys

import numpy as np
import pytest

import libint_amd as la
import oracle
from libint_amd import molecule
from tests.helpers import compare_arbitrated

pytestmark = pytest.mark.gpu

THR = 1e-6
EPS32 = float(np.finfo(np.float32).eps)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def fp32_everywhere(ctx):
    """The FP32 variants of every class (by default mixed batches use only
    those measured faster than FP64)."""
    ctx.set_fp32_policy(la.FP32_ALWAYS)
    yield
    ctx.set_fp32_policy(la.FP32_MEASURED)


def test_mixed_default_routes_only_measured_classes(ctx):
    """With the default routing a mixed plan sends quartets to FP32 only in
    the classes whose FP32 variant the autotuner measured faster
    (fp32_paying_classes); every other class gives the FP64 plan's values
    bit for bit."""
    ctx.set_fp32_policy(la.FP32_MEASURED)
    mol = molecule.make_molecule(2, natoms=4)
    b = la.Basis(ctx, mol.basis)
    q = molecule.unique_quartets(mol.basis.nshell)[:20000]
    say(f"case: mixed plan, default FP32 routing, {len(q)} quartets")
    m = la.EriPlan(b, q, fp32_threshold=THR)
    m.execute()
    gm, _ = m.download()
    p = la.EriPlan(b, q)
    p.execute()
    g, _ = p.download()
    paying = set(la.fp32_paying_classes())
    st = m.stats()
    say(f"  classes whose FP32 variant pays: {sorted(paying)}; FP32 in the plan: "
        f"{sorted((s['cls'], s['nquartets']) for s in st if s['fp32'])}")
    assert all(s["cls"] in paying for s in st if s["fp32"])
    # quartets of the other classes: the FP64 plan's bits
    B = mol.basis
    ls = B.l[q]
    bra = np.sort(ls[:, :2], axis=1)[:, ::-1]
    ket = np.sort(ls[:, 2:], axis=1)[:, ::-1]
    swap = bra.sum(1) > ket.sum(1)
    c1 = np.where(swap[:, None], ket, bra)
    c2 = np.where(swap[:, None], bra, ket)
    cls = np.array([f"{a}{b_}{c}{d}" for (a, b_), (c, d) in zip(c1, c2)])
    sizes = B.sizes()
    n = sizes[q[:, 0]] * sizes[q[:, 1]] * sizes[q[:, 2]] * sizes[q[:, 3]]
    other = np.repeat(~np.isin(cls, list(paying)), n)
    assert other.any()
    assert np.array_equal(gm[other], g[other])
    # a batch large enough to fill the chip with FP32 work: the (dd|dd)
    # list of the 60-atom workload
    mol = molecule.make_molecule(2)
    b = la.Basis(ctx, mol.basis)
    q, _ = molecule.headline_quartets(b)
    say(f"case: mixed plan, default FP32 routing, the 60-atom (dd|dd) list ({len(q)} quartets)")
    m = la.EriPlan(b, q, fp32_threshold=THR)
    st = m.stats()
    n32 = sum(s["nquartets"] for s in st if s["fp32"])
    say(f"  {n32} of {len(q)} quartets in FP32")
    if "2222" in paying:
        assert n32 > 0
    m.execute()
    gm, _ = m.download()
    p = la.EriPlan(b, q)
    p.execute()
    g, _ = p.download()
    assert np.abs(gm - g).max() <= 1e-5 * THR


@pytest.mark.parametrize("workload,natoms,nq", [(2, 6, 60000), (3, 3, 12000)])
def test_mixed_eri_per_class_errors(ctx, workload, natoms, nq):
    """Workload 2 (s-d) and the high-angular-momentum recipe (s-f, K <= 8,
    at workload 3's atom density)."""
    mol = molecule.make_molecule(workload, natoms=natoms, same_density=workload == 3)
    B = mol.basis
    b = la.Basis(ctx, B)
    allq = molecule.unique_quartets(B.nshell)
    rng = np.random.default_rng(4)
    q = allq[rng.choice(len(allq), min(nq, len(allq)), replace=False)]
    say(f"case: mixed ERI, {len(q)} quartets of the workload-{workload} recipe ({natoms} atoms)")
    plan = la.EriPlan(b, q, fp32_threshold=THR)
    offs = plan.offsets()
    plan.execute()
    g, _ = plan.download()
    st = plan.stats()
    n32 = sum(s["nquartets"] for s in st if s["fp32"])
    o, ooff, _ = oracle.eri_quartets(B, q)
    assert np.array_equal(offs, ooff)
    # per-pair K and S: a quartet runs in FP32 iff K K < THR and S S < 16 THR
    i, j = np.tril_indices(B.nshell)
    Kv, Sv = la.precision_bounds(b, np.stack([i, j], 1))
    K = np.zeros((B.nshell, B.nshell))
    S = np.zeros_like(K)
    K[i, j] = K[j, i] = Kv
    S[i, j] = S[j, i] = Sv
    KK = K[q[:, 0], q[:, 1]] * K[q[:, 2], q[:, 3]]
    SS = S[q[:, 0], q[:, 1]] * S[q[:, 2], q[:, 3]]
    # quartets with a shell above the FP32 variants' l run in FP64 regardless
    sel = (KK < THR) & (SS < la.FP32_HRR_FACTOR * THR) & (B.l[q].max(axis=1) <= la.fp32_max_l())
    sizes = B.sizes()
    n = sizes[q[:, 0]] * sizes[q[:, 1]] * sizes[q[:, 2]] * sizes[q[:, 3]]
    is32 = np.repeat(sel, n)
    say(f"  {n32} of {len(q)} quartets in FP32 ({int((KK < THR).sum())} below the Schwarz threshold)")
    # per-class errors of every class, FP32 and FP64 parts
    ends0 = offs + n
    cls0 = np.array(["".join(map(str, r)) for r in B.l[q]])
    for c in sorted(set(cls0)):
        for part, msk in (("fp64", (cls0 == c) & ~sel), ("fp32", (cls0 == c) & sel)):
            if not msk.any():
                continue
            gm = np.concatenate([g[a:e] for a, e in zip(offs[msk], ends0[msk])])
            om = np.concatenate([o[a:e] for a, e in zip(offs[msk], ends0[msk])])
            big = np.abs(om) > 1e-14
            rel = (np.abs(gm - om)[big] / np.abs(om[big])).max(initial=0)
            say(f"  class ({c[:2]}|{c[2:]}) {part} x{int(msk.sum())}: max abs {np.abs(gm - om).max():.2e}, "
                f"max rel (|ref|>1e-14) {rel:.2e}")
    assert n32 > 0 and n32 == int(sel.sum())
    # FP64 quartets at the integral bar, oracle misses settled by brute force
    q64 = ~sel
    g64 = np.concatenate([g[a:a + m] for a, m in zip(offs[q64], n[q64])])
    o64 = np.concatenate([o[a:a + m] for a, m in zip(offs[q64], n[q64])])
    off64 = np.concatenate([[0], np.cumsum(n[q64])[:-1]])
    nf64, narb, oerr = compare_arbitrated(B, q[q64], off64, g64, o64, 1e-12, 1e-10)
    say(f"  FP64 part: fails {nf64}; {narb} values arbitrated by brute force "
        f"(oracle's own error there up to {oerr:.2e})")
    # FP32 quartets: error per quartet against eps_fp32 S_ab S_cd
    ends = offs + n
    e32 = np.array([np.abs(g[a:e] - o[a:e]).max() for a, e in zip(offs[sel], ends[sel])])
    ratio = e32 / np.maximum(EPS32 * SS[sel], 1e-300)
    say(f"  FP32 part: max abs {e32.max():.3e}; max err/(eps32 S S) {ratio[e32 > 1e-16].max(initial=0):.2f}")
    assert nf64 == 0
    # FP32 bar: 1e-5 of the threshold absolute, and within 64 eps_fp32 S S
    assert e32.max() <= 1e-5 * THR
    assert (e32 <= 64 * EPS32 * SS[sel] + 1e-16).all()


def test_mixed_fock_error(ctx):
    mol = molecule.make_molecule(2, natoms=4)
    rng = np.random.default_rng(12)
    C = rng.normal(0.0, 0.1, (mol.nbf, mol.nbf // 5))
    D = C @ C.T
    b = la.Basis(ctx, mol.basis)
    fb = la.FockBuilder(b)
    G64, J64, K64 = fb.build(D, 1e-12)
    fb.set_fp32_threshold(THR)
    G32, J32, K32 = fb.build(D, 1e-12)
    st = fb.stats()
    Go, Jo, Ko, _ = oracle.fock_2e(mol.basis, D, 1e-12)
    say(f"case: mixed Fock nbf={mol.nbf}: {st['nquartets_fp32']} of {st['nquartets']} quartets FP32; "
        f"|G32-Go| {np.abs(G32 - Go).max():.2e} |G64-Go| {np.abs(G64 - Go).max():.2e}")
    assert st["nquartets_fp32"] > 0
    assert np.abs(G32 - Go).max() <= 1e-8
    assert np.abs(G64 - Go).max() <= 1e-8
