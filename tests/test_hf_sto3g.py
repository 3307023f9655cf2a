"""End-to-end closed-shell Hartree-Fock of the reference's water test
(tests/hartree-fock/hartree-fock.cc with h2o.xyz and its built-in STO-3G,
:338-420): the energy must equal the reference's pinned
-74.942080057696 within its own tolerance 1e-11
(tests/hartree-fock/hartree-fock-validate.py:13-14).  On the CPU every
integral comes from the oracle (one-body restatement pinned here by the
reference's python goldens, python/tests/test_libint2.py:32-36); the GPU test
takes the one-body matrices from l2g_onebody and G = 2J - K from the GPU
Fock build."""
import numpy as np
import pytest

import libint_amd as la
import oracle
from oracle import onebody
from tests.helpers import rhf

EREF = -74.942080057696
ANG2BOHR = 1 / 0.52917721092  # the driver's CODATA 2010 value (hartree-fock.cc:305-306)
XYZ = [(8, (0.00000, -0.07579, 0.00000)), (1, (0.86681, 0.60144, 0.00000)), (1, (-0.86681, 0.60144, 0.00000))]
# STO-3G of the driver (hartree-fock.cc:338-420): (l, exponents, coefficients) per atom
STO3G = {
    1: [(0, [3.425250910, 0.623913730, 0.168855400], [0.15432897, 0.53532814, 0.44463454])],
    8: [(0, [130.709320000, 23.808861000, 6.443608300], [0.15432897, 0.53532814, 0.44463454]),
        (0, [5.033151300, 1.169596100, 0.380389000], [-0.09996723, 0.39951283, 0.70011547]),
        (1, [5.033151300, 1.169596100, 0.380389000], [0.15591627, 0.60768372, 0.39195739])],
}


def water():
    atoms = [(Z, np.array(x) * ANG2BOHR) for Z, x in XYZ]
    shells = [la.Shell(l, list(zip(a, c)), tuple(pos), pure=False) for Z, pos in atoms for (l, a, c) in STO3G[Z]]
    enuc = sum(atoms[i][0] * atoms[j][0] / np.linalg.norm(atoms[i][1] - atoms[j][1])
               for i in range(3) for j in range(i))
    return atoms, shells, enuc


def one_body_matrices(atoms, shells):
    sh = [(s.l, s.alpha, s.coeff, s.O) for s in shells]
    S, T, V = onebody.one_body(sh, [(float(Z), pos) for Z, pos in atoms])
    return S, T + V


def test_one_body_goldens():
    """python/tests/test_libint2.py:32-36 (positions used as given, in bohr)."""
    s = la.Shell(0, [(1.0, 10.0)])
    p = la.Shell(1, [(1.0, 10.0)])
    h2o = [(8.0, np.array([0.0, -0.07579, 0.0])), (1.0, np.array([0.86681, 0.60144, 0.0])),
           (1.0, np.array([-0.86681, 0.60144, 0.0]))]
    S, T, V = onebody.one_body([(0, s.alpha, s.coeff, s.O)], h2o)
    assert abs(np.linalg.norm(T) - 1.5) < 1e-12
    assert abs(np.linalg.norm(S) - 1.0) < 1e-12
    assert abs(np.linalg.norm(V) - 14.54704336519) < 1e-10
    sh = [(x.l, x.alpha, x.coeff, x.O) for x in (s, p, s, p)]
    S4, _, _ = onebody.one_body(sh, [])
    assert abs(S4.sum() - 16.0) < 1e-12


def test_rhf_energy_oracle():
    atoms, shells, enuc = water()
    S, H = one_body_matrices(atoms, shells)
    arrays = la.BasisArrays.from_shells(shells)

    def g(D):
        G, _, _, _ = oracle.fock_2e(arrays, D, precision=1e-15)
        return G

    e = rhf(S, H, 5, g, enuc)
    assert abs(e - EREF) < 1e-11, e


@pytest.mark.gpu
def test_rhf_energy_gpu(ctx):
    """Every integral on the GPU: one-body matrices from l2g_onebody, G from
    the Fock build."""
    atoms, shells, enuc = water()
    b = la.Basis(ctx, la.BasisArrays.from_shells(shells))
    S, T, V = la.one_body(b, [(float(Z), pos) for Z, pos in atoms])
    H = T + V
    fb = la.FockBuilder(b)

    def g(D):
        G, _, _ = fb.build(D, 1e-15)
        return G

    e = rhf(S, H, 5, g, enuc)
    print(f"RHF/STO-3G water on the GPU Fock build: {e:.12f} (reference {EREF})", flush=True)
    assert abs(e - EREF) < 1e-11, e
