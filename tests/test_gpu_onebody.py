"""One-body integrals on the GPU (l2g_onebody): overlap, kinetic energy and
point-charge attraction against the oracle's McMurchie-Davidson restatement
(oracle/onebody.py) at the integral bar, on random contracted shells to l = 3
with mixed purity; the reference's python goldens
(python/tests/test_libint2.py:32-36); the Gaussian solid-harmonics ordering
as a permutation of the Standard one."""
import sys

import numpy as np
import pytest

import libint_amd as la
from oracle import onebody
from tests.helpers import cart_to_shells, compare, random_shells

pytestmark = pytest.mark.gpu

ABS, REL = 1e-12, 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_one_body(shells, charges):
    S, T, V = onebody.one_body([(s.l, s.alpha, s.coeff, s.O) for s in shells], charges)
    X = cart_to_shells(shells)
    return X.T @ S @ X, X.T @ T @ X, X.T @ V @ X


def test_one_body_vs_oracle(ctx):
    rng = np.random.default_rng(91)
    shells = random_shells(rng, [0, 1, 2, 3, 1, 2, 3, 0], kmax=3)
    shells[5].pure = False
    shells[6].pure = False
    charges = [(8.0, rng.normal(0, 1.2, 3)), (1.0, rng.normal(0, 1.2, 3)), (2.5, rng.normal(0, 1.2, 3))]
    say(f"case: one-body S, T, V, {len(shells)} shells to l = 3, 3 charges")
    S, T, V = la.one_body(la.Basis(ctx, la.BasisArrays.from_shells(shells)), charges)
    So, To, Vo = oracle_one_body(shells, charges)
    for name, g, o in (("S", S, So), ("T", T, To), ("V", V, Vo)):
        nf = compare(g.reshape(-1), o.reshape(-1), ABS, REL)[0]
        say(f"  {name}: max abs {np.abs(g - o).max():.2e}, fails {nf}")
        assert nf == 0, name
    assert np.allclose(S, S.T, atol=1e-14) and np.allclose(V, V.T, atol=1e-13)


def test_one_body_goldens(ctx):
    """python/tests/test_libint2.py:32-36 (positions as given, in bohr)."""
    s = la.Shell(0, [(1.0, 10.0)])
    p = la.Shell(1, [(1.0, 10.0)])
    h2o = [(8.0, (0.0, -0.07579, 0.0)), (1.0, (0.86681, 0.60144, 0.0)), (1.0, (-0.86681, 0.60144, 0.0))]
    say("case: one-body python goldens")
    S, T, V = la.one_body(la.Basis(ctx, la.BasisArrays.from_shells([s])), h2o)
    assert abs(np.linalg.norm(T) - 1.5) < 1e-12
    assert abs(np.linalg.norm(S) - 1.0) < 1e-12
    assert abs(np.linalg.norm(V) - 14.54704336519) < 1e-10
    S4, _, _ = la.one_body(la.Basis(ctx, la.BasisArrays.from_shells([s, p, s, p])), [])
    assert abs(S4.sum() - 16.0) < 1e-12


def test_one_body_gaussian_ordering(ctx):
    rng = np.random.default_rng(92)
    shells = random_shells(rng, [2, 3, 1], kmax=2)
    b = la.Basis(ctx, la.BasisArrays.from_shells(shells))
    charges = [(3.0, (0.1, 0.2, 0.3))]
    say("case: one-body Gaussian ordering")
    S, T, V = la.one_body(b, charges)
    ctx.set_solid_harmonics_ordering(la.SHG_GAUSSIAN)
    try:
        Sg, Tg, Vg = la.one_body(b, charges)
    finally:
        ctx.set_solid_harmonics_ordering(la.SHG_STANDARD)
    perm = []
    off = 0
    for s in shells:
        n = 2 * s.l + 1 if (s.pure and s.l > 0) else (s.l + 1) * (s.l + 2) // 2
        if s.pure and s.l > 0:
            perm += [off + s.l + ((q + 1) // 2 if q % 2 else -(q // 2)) for q in range(n)]
        else:
            perm += list(range(off, off + n))
        off += n
    for a, g in ((S, Sg), (T, Tg), (V, Vg)):
        assert np.array_equal(a[np.ix_(perm, perm)], g)


def test_engine_one_body_and_erf(ctx):
    """The Engine path of the reference's python binding: one-body operators
    with BraKet XX and the charges through set_params
    (python/tests/test_libint2.py:32-36 goldens), and erf_coulomb with omega
    through set_params, against the plan path."""
    say("case: Engine one-body operators and erf_coulomb")
    s = la.Shell(0, [(1.0, 10.0)])
    h2o = [(8.0, (0.0, -0.07579, 0.0)), (1.0, (0.86681, 0.60144, 0.0)), (1.0, (-0.86681, 0.60144, 0.0))]
    assert abs(np.linalg.norm(la.Engine(1, 0, ctx=ctx, oper=la.OVERLAP).compute(s, s)) - 1.0) < 1e-12
    assert abs(np.linalg.norm(la.Engine(1, 0, ctx=ctx, oper=la.KINETIC).compute(s, s)) - 1.5) < 1e-12
    en = la.Engine(1, 0, ctx=ctx, oper=la.NUCLEAR)
    with pytest.raises(ValueError):
        en.compute(s, s)
    assert abs(np.linalg.norm(en.set_params(h2o).compute(s, s)) - 14.54704336519) < 1e-10
    rng = np.random.default_rng(93)
    sh = random_shells(rng, [2, 1, 0, 2], kmax=2)
    v = la.Engine(2, 2, ctx=ctx, oper=la.ERF_COULOMB).set_params(0.4).compute(*sh)
    plan = la.EriPlan(la.Basis(ctx, la.BasisArrays.from_shells(sh)), [[0, 1, 2, 3]])
    plan.set_operator(la.ERF_COULOMB, 0.4)
    plan.execute()
    w, _ = plan.download()
    assert np.array_equal(v, w)
    with pytest.raises(ValueError):
        la.Engine(1, 0, ctx=ctx, oper=la.OVERLAP, deriv_order=1)


def test_engine_one_body_prescaled_uniform(ctx):
    """prescale_by and uniform Cartesian normalization on the one-body path:
    the Engine's block is the oracle's times the scale and the per-function
    uniform factors (engine.impl.h:1050, 552-558); deriv_order 1 refuses both,
    as the C++ façade does."""
    say("case: Engine one-body prescaled, uniform Cartesian normalization")
    rng = np.random.default_rng(94)
    sh = random_shells(rng, [2, 1, 3], kmax=2, pure=False)
    S, T, V = oracle_one_body(sh, [(3.0, np.array([0.2, -0.1, 0.4]))])
    off = np.concatenate([[0], np.cumsum([s.size() for s in sh])])
    for oper, M in ((la.OVERLAP, S), (la.KINETIC, T), (la.NUCLEAR, V)):
        e = la.Engine(2, 3, ctx=ctx, oper=oper).prescale_by(0.6).set_cartesian_normalization(True)
        if oper == la.NUCLEAR:
            e.set_params([(3.0, (0.2, -0.1, 0.4))])
        for a, b in ((0, 1), (2, 0), (1, 2)):
            ref = 0.6 * M[off[a]:off[a + 1], off[b]:off[b + 1]] * np.outer(la.uniform_cart_scales(sh[a]),
                                                                           la.uniform_cart_scales(sh[b]))
            assert compare(e.compute(sh[a], sh[b]), ref.reshape(-1), ABS, REL)[0] == 0
    ed = la.Engine(2, 2, ctx=ctx, deriv_order=1).prescale_by(2.0)
    with pytest.raises(ValueError):
        ed.compute(*random_shells(rng, [1, 1, 0, 0], kmax=1))
