"""The C++ façade (include/libint2_amd/engine.h): compiles with g++ against
the C ABI, its CPU-side error behaviour holds, and (GPU) its results equal
the oracle's for the same shells."""
import os
import subprocess

import numpy as np
import pytest

import libint_amd as la
import oracle
import oracle.deriv1
from oracle import onebody
from tests.helpers import cart_to_shells, compare

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "facade_test")


def build_exe():
    src = os.path.join(ROOT, "tests", "cpp", "facade_test.cpp")
    libdir = os.path.dirname(la.LIBPATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-o", EXE,
                    "-L", libdir, "-ll2g", f"-Wl,-rpath,{libdir}"], check=True)


def test_facade_compiles_and_cpu_checks():
    build_exe()
    r = subprocess.run([EXE, "cpu"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_facade_results_match_oracle():
    if not os.path.exists(EXE):
        build_exe()
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    shells = [
        (0, [0.9, 3.1], [0.4, 0.7], (0.1, 0.2, 0.3), False),
        (1, [1.3, 0.4], [0.5, 0.6], (1.0, 0.0, -0.2), False),
        (2, [2.2], [1.0], (-0.3, 0.8, 0.1), True),
        (2, [0.7, 1.9], [0.3, 0.9], (0.4, -0.5, 0.6), False),
    ]
    sh = [la.Shell(l, list(zip(a, c)), o, pure=p) for (l, a, c, o, p) in shells] + [la.Shell.unit()]
    arrays = la.BasisArrays.from_shells(sh)
    lines = r.stdout.split("\n")
    i = 0
    checked = 0
    while i < len(lines):
        if lines[i].startswith("D "):
            # first derivatives: 12 shell sets of n, against the oracle's restatement
            _, a, b, c, d, n = lines[i].split()
            n = int(n)
            vals = np.array([float(x) for x in lines[i + 1:i + 1 + 12 * n]]).reshape(12, n)
            o, _, _ = oracle.deriv1.eri_deriv1(arrays, [[int(a), int(b), int(c), int(d)]])
            for k in range(12):
                assert compare(vals[k], o[k], 1e-12, 1e-10)[0] == 0, k
            checked += 1
            i += 12 * n + 1
        elif lines[i][:2] in ("O ", "K ", "N ", "U "):
            # one-body <a|O|b> block against the oracle's (Cartesian, then the shells' functions)
            kind, a, b, n = lines[i].split()
            n = int(n)
            vals = np.array([float(x) for x in lines[i + 1:i + 1 + n]])
            four = sh[:4]
            charges = [(8.0, np.array([0.1, 0.0, -0.2])), (1.0, np.array([1.0, 0.5, 0.3]))]
            S, T, V = onebody.one_body([(x.l, x.alpha, x.coeff, x.O) for x in four], charges)
            X = cart_to_shells(four)
            M = X.T @ {"O": S, "K": T, "N": V, "U": S}[kind] @ X
            off = np.concatenate([[0], np.cumsum([x.size() for x in four])])
            a, b = int(a), int(b)
            blk = M[off[a]:off[a + 1], off[b]:off[b + 1]]
            if kind == "U":  # prescale_by(1.7), uniform Cartesian normalization
                blk = 1.7 * blk * np.outer(la.uniform_cart_scales(four[a]), la.uniform_cart_scales(four[b]))
            blk = blk.reshape(-1)
            assert compare(vals, blk, 1e-12, 1e-10)[0] == 0, kind
            checked += 1
            i += n + 1
        elif lines[i].startswith("F "):
            # the hartree-fock++ Fock loop through compute2 with SchwarzInf
            # shell pairs, against the oracle's compute_2body_fock
            nbf = int(lines[i].split()[1])
            vals = np.array([float(x) for x in lines[i + 1:i + 1 + 2 * nbf * nbf]])
            D, G = vals[:nbf * nbf].reshape(nbf, nbf), vals[nbf * nbf:].reshape(nbf, nbf)
            fb = []
            for o in ((0.0, 0.0, 0.0), (1.4, 0.3, -0.2), (-0.5, 1.6, 0.4)):
                fb += [la.Shell(0, list(zip([5.0, 1.2, 0.3], [0.3, 0.5, 0.4])), o, pure=False),
                       la.Shell(1, list(zip([2.1, 0.45], [0.6, 0.5])), o, pure=False),
                       la.Shell(2, [(0.8, 1.0)], o, pure=True)]
            Go, _, _, _ = oracle.fock_2e(la.BasisArrays.from_shells(fb), D, 1e-12)
            print(f"facade Fock loop (compute2 with SchwarzInf ShellPairs), nbf {nbf}: "
                  f"max |G - G_oracle| {np.abs(G - Go).max():.2e}")
            assert np.abs(G - Go).max() <= 1e-10
            checked += 1
            i += 2 * nbf * nbf + 1
        elif lines[i].startswith("LAT "):
            print(f"facade Engine::compute latency, contracted (dd|dd): {float(lines[i].split()[1]):.1f} us per call")
            i += 1
        elif lines[i].startswith("Q ") or lines[i].startswith("R "):
            kind, a, b, c, d, n = lines[i].split()
            n = int(n)
            vals = np.array([float(x) for x in lines[i + 1:i + 1 + n]])
            # R: erf_coulomb with omega 0.7 (the façade's Engine params)
            op = dict(oper=oracle.ERF_COULOMB, omega=0.7) if kind == "R" else {}
            o, _, _ = oracle.eri_quartets(arrays, [[int(a), int(b), int(c), int(d)]], **op)
            assert compare(vals, o, 1e-12, 1e-10)[0] == 0
            checked += 1
            i += n + 1
        else:
            i += 1
    # 5 quartets, one 3-centre, one 2-centre, one erf set, one derivative set, four one-body blocks,
    # the Fock loop
    assert checked == 14
