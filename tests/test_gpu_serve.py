"""Fock-build request server: continuous batching of requests for different
molecules.  Each result must equal the same request run alone (1e-12
absolute: only the order of the FP64 atomic additions differs) and the
oracle's compute_2body_fock restatement (1e-8, as the single-build tests)."""
import sys
import time

import numpy as np
import pytest

import libint_amd as la
import oracle
from libint_amd import molecule
from tests.helpers import random_shells

pytestmark = pytest.mark.gpu

ALONE_TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def density(nbf, seed):
    rng = np.random.default_rng(seed)
    C = rng.normal(0.0, 0.1, (nbf, max(1, nbf // 5)))
    return C @ C.T


def requests():
    out = []
    for k, (w, n) in enumerate([(1, 4), (2, 3), (1, 6), (2, 2), (1, 3), (2, 4)]):
        mol = molecule.make_molecule(w, natoms=n)
        out.append((mol.basis, density(mol.nbf, 100 + k)))
    return out


def test_serve_alone_equals_mix():
    reqs = requests()
    say(f"case: server alone vs mix, {len(reqs)} requests, nbf {[D.shape[0] for _, D in reqs]}")
    srv = la.Server(0, round_quartets=256)
    try:
        alone = []
        for B, D in reqs:
            alone.append(srv.result(srv.submit(B, D, 1e-12)))
        st0 = srv.stats()
        ids = [srv.submit(B, D, 1e-12) for B, D in reqs]
        mix = [srv.result(i) for i in ids]
        st1 = srv.stats()
    finally:
        srv.close()
    say(f"  rounds alone {st0['rounds']}, mix {st1['rounds'] - st0['rounds']}; "
        f"most requests in one round {st1['max_requests_per_round']}")
    assert st1["max_requests_per_round"] >= 2
    for (B, D), a, m in zip(reqs, alone, mix):
        for x, y in zip(a[:3], m[:3]):
            assert np.abs(x - y).max() <= ALONE_TOL
    # the first two against the oracle
    for (B, D), a in list(zip(reqs, alone))[:2]:
        Go, Jo, Ko, _ = oracle.fock_2e(B, D, 1e-12)
        say(f"  vs oracle: max|G-Go| {np.abs(a[0] - Go).max():.2e}")
        assert np.abs(a[0] - Go).max() <= 1e-8
        assert np.abs(a[1] - Jo).max() <= 1e-8
        assert np.abs(a[2] - Ko).max() <= 1e-8


def test_serve_deterministic_alone_equals_mix_bitwise():
    """With the deterministic digest a request's J, K and G have the same
    bits alone and in a mix with others."""
    reqs = requests()[:4]
    say(f"case: deterministic server alone vs mix, {len(reqs)} requests")
    srv = la.Server(0, round_quartets=256, deterministic=True)
    try:
        alone = [srv.result(srv.submit(B, D, 1e-12)) for B, D in reqs]
        ids = [srv.submit(B, D, 1e-12) for B, D in reqs]
        mix = [srv.result(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert st["max_requests_per_round"] >= 2
    for a, m in zip(alone, mix):
        for x, y in zip(a[:3], m[:3]):
            assert np.array_equal(x, y)
    Go, _, _, _ = oracle.fock_2e(reqs[0][0], reqs[0][1], 1e-12)
    assert np.abs(alone[0][0] - Go).max() <= 1e-8


def test_serve_equals_single_build(ctx):
    mol = molecule.make_molecule(2, natoms=3)
    D = density(mol.nbf, 7)
    b = la.Basis(ctx, mol.basis)
    fb = la.FockBuilder(b)
    G1, J1, K1 = fb.build(D, 1e-12)
    say("case: server vs FockBuilder, same molecule")
    srv = la.Server(0)
    try:
        G2, J2, K2, ms = srv.result(srv.submit(mol.basis, D, 1e-12))
    finally:
        srv.close()
    assert np.abs(G1 - G2).max() <= ALONE_TOL
    assert ms[1] >= ms[0] >= 0


def test_serve_memory_admission_and_failures():
    """A budget that holds one request at a time still completes them all in
    order; a request larger than the whole budget fails with L2G_E_OOM and a
    request above the compiled l fails with L2G_E_LMAX, neither disturbing
    the others."""
    # uncontracted shells on close centres: every shell pair and primitive
    # pair is kept, so the matrix arena (D, jg, kg: 24 bytes per element,
    # 60% of the budget) binds before the primitive-pair (88 bytes, 35%) and
    # shell-pair (32 bytes, 5%) arenas; the budget holds 1.5 requests
    rng = np.random.default_rng(21)
    shells = random_shells(rng, [2, 2, 2, 1, 1, 0, 2, 1], kmax=1, box=(0.0, 1.5), expo=(0.5, 2.0))
    arrays = la.BasisArrays.from_shells(shells)
    nbf = int(arrays.sizes().sum())
    npairs = arrays.nshell * (arrays.nshell + 1) // 2
    budget = 1.5 * 24 * nbf * nbf / 0.60
    assert 0.35 * budget / 88 >= npairs and 0.05 * budget / 32 >= npairs
    small = [arrays] * 3
    big = molecule.make_molecule(2, natoms=12)
    say(f"case: server admission by memory, budget {budget / 1e6:.3f} MB, nbf {nbf}")
    srv = la.Server(0, hbm_bytes=budget)
    try:
        ids = [srv.submit(a, density(nbf, 5 + k), 1e-12) for k, a in enumerate(small)]
        bad = srv.submit(big.basis, density(big.nbf, 1), 1e-12)
        res = [srv.result(i) for i in ids]
        with pytest.raises(la.L2gError) as e:
            srv.wait(bad)
        assert e.value.code == la.L2G_E_OOM
        st = srv.stats()
        assert st["max_requests_per_round"] == 1
        fb = la.BasisArrays.from_shells([la.Shell(la.max_l() + 1, [(1.0, 1.0)])])
        nf = int(fb.sizes().sum())
        with pytest.raises(la.L2gError) as e2:
            srv.wait(srv.submit(fb, np.eye(nf), 1e-12))
        assert e2.value.code == la.L2G_E_LMAX
        # the server keeps working after failures
        G, _, _, _ = srv.result(srv.submit(small[0], density(nbf, 5), 1e-12))
        assert np.abs(G - res[0][0]).max() <= ALONE_TOL
    finally:
        srv.close()


def test_serve_latency_under_load():
    """Poisson arrivals of requests for molecules of 2-8 atoms; every request
    completes and the p50/p99 latencies are reported."""
    rng = np.random.default_rng(9)
    mols = [molecule.make_molecule(1 + (k % 2), natoms=int(rng.integers(2, 9))) for k in range(24)]
    dens = [density(m.nbf, 200 + k) for k, m in enumerate(mols)]
    say(f"case: server under load, {len(mols)} requests")
    srv = la.Server(0, round_quartets=1 << 16)
    try:
        ids = []
        t0 = time.perf_counter()
        for m, D in zip(mols, dens):
            time.sleep(rng.exponential(0.005))
            ids.append(srv.submit(m.basis, D, 1e-12))
        lat = np.array([srv.result(i)[3][1] for i in ids])
        wall = time.perf_counter() - t0
        st = srv.stats()
    finally:
        srv.close()
    say(f"  {len(ids)} requests in {wall:.2f}s; latency p50 {np.percentile(lat, 50):.1f} ms "
        f"p99 {np.percentile(lat, 99):.1f} ms; rounds {st['rounds']}, most requests in a round "
        f"{st['max_requests_per_round']}")
    assert len(lat) == len(mols) and (lat > 0).all()
