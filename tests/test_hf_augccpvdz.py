"""End-to-end closed-shell Hartree-Fock of the reference's second water test
(tests/hartree-fock/hartree-fock++.cc with h2o_rotated.xyz and aug-cc-pVDZ,
pure d shells): the energy must equal the reference's pinned
-76.003354058439 within its tolerance 5e-12
(tests/hartree-fock/hartree-fock++-validate.py:47-48), with G = 2J - K from
the GPU Fock build.  Basis and geometry: tests/golden/h2o_augccpvdz.json
(made by tests/golden/make_h2o_augccpvdz.py from the reference's data files);
one-body integrals from the oracle's restatement."""
import json
import os

import numpy as np
import pytest

import libint_amd as la
from oracle import onebody
from tests.helpers import rhf

HERE = os.path.dirname(os.path.abspath(__file__))
BOHR_TO_ANGSTROM = 0.529177210903  # CODATA 2018, read_dotxyz's default (atom.h:50-55, 209-216)
ZOF = {"H": 1, "O": 8}


def water():
    d = json.load(open(os.path.join(HERE, "golden", "h2o_augccpvdz.json")))
    atoms = [(ZOF[el], np.array(x) / BOHR_TO_ANGSTROM) for el, x in d["atoms_angstrom"]]
    shells = []
    for (Z, pos), (el, _) in zip(atoms, d["atoms_angstrom"]):
        for s in d["basis"][el]:
            shells.append(la.Shell(s["l"], list(zip(s["exponents"], s["coefficients"])), tuple(pos),
                                   pure=s["l"] > 1))
    enuc = sum(atoms[i][0] * atoms[j][0] / np.linalg.norm(atoms[i][1] - atoms[j][1])
               for i in range(len(atoms)) for j in range(i))
    return atoms, shells, enuc, d["energy"], d["energy_tolerance"]


def one_body_pure(atoms, shells):
    """S and H = T + V in the shells' own (pure or Cartesian) functions."""
    S, T, V = onebody.one_body([(s.l, s.alpha, s.coeff, s.O) for s in shells],
                               [(float(Z), pos) for Z, pos in atoms])
    blocks = []
    for s in shells:
        nc = (s.l + 1) * (s.l + 2) // 2
        if s.pure and s.l > 0:
            blocks.append(np.array([[la.solid_harmonics_coeff(s.l, m, *c) for c in onebody.cart(s.l)]
                                    for m in range(-s.l, s.l + 1)]).T)
        else:
            blocks.append(np.eye(nc))
    nct = sum(b.shape[0] for b in blocks)
    nft = sum(b.shape[1] for b in blocks)
    Tr = np.zeros((nct, nft))
    r = c = 0
    for b in blocks:
        Tr[r:r + b.shape[0], c:c + b.shape[1]] = b
        r += b.shape[0]
        c += b.shape[1]
    return Tr.T @ S @ Tr, Tr.T @ (T + V) @ Tr


@pytest.mark.gpu
def test_rhf_energy_gpu_augccpvdz(ctx):
    atoms, shells, enuc, eref, tol = water()
    S, H = one_body_pure(atoms, shells)
    fb = la.FockBuilder(la.Basis(ctx, la.BasisArrays.from_shells(shells)))

    def g(D):
        G, _, _ = fb.build(D, 1e-15)
        return G

    e = rhf(S, H, 5, g, enuc, diis=8)
    print(f"RHF/aug-cc-pVDZ rotated water on the GPU Fock build: {e:.12f} (reference {eref})", flush=True)
    assert abs(e - eref) < tol, e


def test_rhf_energy_oracle_augccpvdz():
    """The same SCF with the oracle's G: pins the one-body restatement, the
    pure-d transform and the basis assembly independently of the GPU."""
    import oracle
    atoms, shells, enuc, eref, tol = water()
    S, H = one_body_pure(atoms, shells)
    arrays = la.BasisArrays.from_shells(shells)

    def g(D):
        G, _, _, _ = oracle.fock_2e(arrays, D, precision=1e-15)
        return G

    e = rhf(S, H, 5, g, enuc, diis=8)
    assert abs(e - eref) < tol, e
