"""The resident single-quartet evaluator (l2g_quartet_engine_*, the per-call
path of Engine::compute / compute2): its results are the batched plan's, bit
for bit, for every class to l = 3 in random caller order, for the three
operators and both solid-harmonic orderings; precomputed primitive-pair data
(compute2 with ShellPair arguments, engine.impl.h:1249-1285) that keep every
primitive pair reproduce the unscreened plan; the Python Engine runs on it
and meets the oracle at the per-integral bar."""
import itertools
import sys

import numpy as np
import pytest

import libint_amd as la
import oracle
from tests.helpers import compare, random_shells

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_quartet_engine_equals_plan(ctx):
    rng = np.random.default_rng(404)
    lmax = la.max_l()
    # K <= 2: no class reaches the plan's few-and-deep routing to the generic kernel
    shells = random_shells(rng, [l for l in range(lmax + 1) for _ in range(2)], kmax=2)
    n = len(shells)
    allq = np.array(list(itertools.product(range(n), repeat=4)), np.int32)
    q = allq[rng.choice(len(allq), 400, replace=False)]
    say(f"case: quartet engine vs plan, {len(q)} quartets to l = {lmax}")
    arrays = la.BasisArrays.from_shells(shells)
    b = la.Basis(ctx, arrays)
    qe = la.QuartetEngine(ctx, 3, lmax)
    for oper, omega in ((la.COULOMB, 0.0), (la.ERF_COULOMB, 0.6), (la.ERFC_COULOMB, 0.6)):
        for order in (la.SHG_STANDARD, la.SHG_GAUSSIAN):
            ctx.set_solid_harmonics_ordering(order)
            try:
                P = la.EriPlan(b, q)
                P.set_wave_split(False)  # a lane per quartet, as the quartet engine runs
                if oper != la.COULOMB:
                    P.set_operator(oper, omega)
                offs = P.offsets()
                P.execute()
                ref, refnz = P.download()
                ends = np.append(offs[1:], ref.size)
                for k, quad in enumerate(q):
                    v, nz = qe.compute([shells[i] for i in quad], oper=oper, omega=omega)
                    assert nz == bool(refnz[k])
                    assert np.array_equal(v, ref[offs[k]:ends[k]]), (oper, order, quad)
                P._release()
            finally:
                ctx.set_solid_harmonics_ordering(la.SHG_STANDARD)
    qe._release()


def test_quartet_engine_given_pairs(ctx):
    """Given primitive-pair data that keep every primitive pair (the pairs'
    own ln_scr, SchwarzInf screening: no per-primitive prefactor screen) give
    the plan's unscreened values; dropping every primitive pair of the bra
    gives the screened-out flag."""
    rng = np.random.default_rng(405)
    shells = random_shells(rng, [2, 1, 2, 0, 3, 1], kmax=2)
    qe = la.QuartetEngine(ctx, 3, 3)
    b = la.Basis(ctx, la.BasisArrays.from_shells(shells))
    q = np.array([[0, 1, 2, 3], [4, 2, 1, 0], [2, 5, 4, 4], [3, 3, 0, 5]], np.int32)
    P = la.EriPlan(b, q, precision=0.0)
    P.set_wave_split(False)
    offs = P.offsets()
    P.execute()
    ref, _ = P.download()
    ends = np.append(offs[1:], ref.size)
    say(f"case: quartet engine with precomputed pair data, {len(q)} quartets")
    for k, quad in enumerate(q):
        s = [shells[i] for i in quad]
        ln = []
        for a, c in ((s[0], s[1]), (s[2], s[3])):
            AB2 = float(((a.O - c.O) ** 2).sum())
            g = a.alpha[:, None] + c.alpha[None, :]
            lns = -(a.alpha[:, None] * c.alpha[None, :] / g) * AB2 + np.log(np.abs(a.coeff))[:, None] + \
                np.log(np.abs(c.coeff))[None, :]
            ln.append(lns.reshape(-1))
        v, nz = qe.compute(s, precision=0.0, screening=la.SCREEN_SCHWARZINF, ln_scr=np.concatenate(ln))
        assert nz and np.array_equal(v, ref[offs[k]:ends[k]])
        dropped = np.concatenate([np.full_like(ln[0], -np.inf), ln[1]])
        v, nz = qe.compute(s, precision=0.0, screening=la.SCREEN_SCHWARZINF, ln_scr=dropped)
        assert not nz and not v.any()
    with pytest.raises(la.L2gError):  # a Schwarz method needs the pairs
        qe.compute([shells[i] for i in q[0]], screening=la.SCREEN_SCHWARZINF)
    qe._release()


def test_engine_per_call_path_vs_oracle(ctx):
    """Engine.compute (Python drop-in) on the resident evaluator against the
    oracle at the per-integral bar, with prescaling."""
    rng = np.random.default_rng(406)
    shells = random_shells(rng, [0, 1, 2, 3, 2, 1], kmax=3)
    arrays = la.BasisArrays.from_shells(shells)
    e = la.Engine(3, 3, ctx=ctx).prescale_by(0.8)
    q = np.array(list(itertools.product(range(len(shells)), repeat=4)), np.int32)[::11]
    say(f"case: Engine.compute per call, {len(q)} quartets vs oracle")
    o, ooff, onz = oracle.eri_quartets(arrays, q)
    ends = np.append(ooff[1:], o.size)
    for k, quad in enumerate(q):
        v = e.compute(*[shells[i] for i in quad])
        assert (v is not None) == bool(onz[k])
        if v is not None:
            assert compare(v, 0.8 * o[ooff[k]:ends[k]], 1e-12, 1e-10)[0] == 0
