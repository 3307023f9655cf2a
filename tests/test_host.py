"""CPU tests of the shipped host side: the C ABI library loads and exports
everything include/l2g.h declares, its host utilities agree with the oracle,
the generator's bookkeeping is complete, the workload generator matches the
stated sizes.  No GPU compute here."""
import os
import re

import numpy as np
import pytest

import libint_amd as la
import oracle
from libint_amd import flop_model, molecule
from libint_amd.codegen import gen_eri

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS = np.finfo(float).eps


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "l2g.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(l2g_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import ctypes
    L = ctypes.CDLL(la.LIBPATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding table covers them all
    assert set(syms) <= set(la.ABI), set(syms) - set(la.ABI)


def test_version_and_lmax():
    assert b"gfx950" in la.lib().l2g_version()
    assert la.max_l() >= 2


def test_init_without_device_fails_loudly():
    import ctypes
    h = ctypes.c_void_p()
    rc = la.lib().l2g_init(0, ctypes.byref(h))
    if rc == 0:
        la.lib().l2g_finalize(h)
        pytest.skip("a GPU is visible here")
    assert rc == -3
    assert b"device" in la.lib().l2g_last_error()


def test_server_without_device_fails_loudly():
    import ctypes
    h = ctypes.c_void_p()
    cfg = la.ServerConfig()
    rc = la.lib().l2g_server_create(ctypes.byref(cfg), ctypes.byref(h))
    if rc == 0:
        la.lib().l2g_server_destroy(h)
        pytest.skip("a GPU is visible here")
    assert rc == la.L2G_E_NODEVICE
    assert not h.value


def test_renorm_and_sph_vs_oracle():
    for l in range(5):
        a = [0.11, 1.7, 30.0]
        c1, m1 = la.shell_renorm(l, a, [0.2, 0.7, 0.4])
        c2, m2 = oracle.shell_renorm(l, a, [0.2, 0.7, 0.4])
        assert np.allclose(c1, c2, rtol=2 * EPS, atol=0)
        assert np.allclose(m1, m2, rtol=0, atol=4 * EPS)
        for m in range(-l, l + 1):
            for lx in range(l + 1):
                for ly in range(l - lx + 1):
                    assert la.solid_harmonics_coeff(l, m, lx, ly, l - lx - ly) == \
                        oracle.solidharm_coeff(l, m, lx, ly, l - lx - ly)
    assert la.shell_renorm(2, [1.0], [1.0])[0][0] == pytest.approx(1.64592278064949, rel=1e-13)


def test_renorm_rejects_bad_input():
    with pytest.raises(la.L2gError):
        la.shell_renorm(1, [-1.0], [1.0])


def test_generator_sph_rows_match_library():
    for l in range(2, 5):
        rows = gen_eri.sh_rows(l)
        for i, m in enumerate(range(-l, l + 1)):
            dense = np.zeros(gen_eri.ncart(l))
            for ic, v in rows[i]:
                dense[ic] = v
            ref = [la.solid_harmonics_coeff(l, m, *c) for c in gen_eri.cart(l)]
            assert np.array_equal(dense, np.array(ref))


def _boys_like_device(tab, T, L):
    """The device algorithm (boys_fm in csrc/eri_device.h) on the host."""
    if T > 117.0:
        f = 0.88622692545275801365 * np.sqrt(1.0 / T)
        out = [f]
        for m in range(1, L + 1):
            f = f * (m - 0.5) * (1.0 / T)
            out.append(f)
        return np.array(out)
    x7 = T * 7.0
    iv = int(x7)
    xd = x7 - iv - 0.5
    c = tab[L, iv]
    f = c[7]
    for k in range(6, -1, -1):
        f = f * xd + c[k]
    fm = np.zeros(L + 1)
    fm[L] = f
    e = np.exp(-T)
    for m in range(L, 0, -1):
        fm[m - 1] = (2 * T * fm[m] + e) * (1.0 / (2 * m - 1))
    return fm


def test_boys_table_and_downward_scheme_vs_mpmath():
    """The library's table (its own construction) read the way the kernels
    read it: F_L interpolated, lower m by downward recursion.  Within the
    reference's Chebyshev tolerance (abs <= eps, rel <= 125 eps)."""
    import json
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "boys_golden.json")))
    T, F = np.array(g["T"]), np.array(g["F"])
    tab = la.boys_table()
    for L in range(0, 41):
        for j, t in enumerate(T):
            v = _boys_like_device(tab, t, L)
            d = np.abs(v - F[j, :L + 1])
            assert (d <= EPS).all(), (L, t)
            assert (d / F[j, :L + 1] <= 125 * EPS).all(), (L, t)


def test_boys_table_vs_oracle_table():
    t = la.boys_table()
    to = oracle.boys_table()
    n = min(t.shape[1], to.shape[0])
    assert np.abs(t[:, :n, :] - np.transpose(to, (1, 0, 2))[:, :n, :]).max() < 4e-15


@pytest.mark.parametrize("budget", [40, 64, 96])
def test_generator_groups_cover_targets_once(budget):
    """The target blocks partition the contracted [e0|f0] set exactly."""
    for c in gen_eri.canonical_classes(2):
        p = gen_eri.plan_class(*c, budget=budget)
        flat = sorted((i, j) for es, fs in p.groups for i in es for j in fs)
        assert flat == [(i, j) for i in range(len(p.ecomps)) for j in range(len(p.fcomps))]


def test_flop_model_matches_canonical_counts():
    # HRR and cart->sph counts of the (dd|dd) canonical model
    assert flop_model.quartet_flops(2, 2, 2, 2) == 11256 + 10736
    # node-by-node and class-by-class counts of the VRR agree
    for cls in [(0, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1), (2, 1, 2, 0), (2, 2, 2, 2), (3, 1, 2, 2), (3, 3, 3, 3)]:
        assert flop_model.vrr_flops(*cls) == flop_model.class_vrr_flops(*cls)
    assert flop_model.vrr_flops(2, 2, 2, 2) == 18261 + 961
    assert flop_model.prim_flops(2, 2, 2, 2) == 204 + 18261 + 961
    assert flop_model.prim_flops(2, 2, 2, 2, model="survey") == 24116
    assert flop_model.prim_flops(2, 1, 2, 2, model="survey") == flop_model.prim_flops(2, 1, 2, 2)
    assert flop_model.prim_flops(0, 0, 0, 0) == 60 + 16 + 1  # prefactors + Boys + 1 contraction add


def test_workload_sizes():
    m1 = molecule.make_molecule(1)
    assert m1.basis.nshell == 60 and m1.nbf == 140
    assert molecule.quartet_count(60) == 1675365
    assert len(molecule.unique_quartets(60)) == 1675365
    assert ((m1.basis.l == 2).sum() == 10)
    m2 = molecule.make_molecule(2)
    assert m2.basis.nshell == 540 and m2.nbf == 1380
    # reproducible
    m1b = molecule.make_molecule(1)
    assert np.array_equal(m1.basis.alpha, m1b.basis.alpha)


def test_unique_quartet_loop_order():
    q = molecule.unique_quartets(4)
    for s1, s2, s3, s4 in q:
        assert s1 >= s2 and s3 <= s1 and s4 <= (s2 if s1 == s3 else s3)
    assert len({tuple(x) for x in q}) == len(q) == molecule.quartet_count(4)
