"""End-to-end closed-shell Hartree-Fock of the reference's second water test
(tests/hartree-fock/hartree-fock++.cc with h2o_rotated.xyz and aug-cc-pVDZ,
pure d shells): the energy must equal the reference's pinned
-76.003354058439 within its tolerance 5e-12
(tests/hartree-fock/hartree-fock++-validate.py:47-48), with G = 2J - K from
the GPU Fock build.  Basis and geometry: tests/golden/h2o_augccpvdz.json
(made by tests/golden/make_h2o_augccpvdz.py from the reference's data files);
one-body integrals from l2g_onebody on the GPU (from the oracle's
restatement in the CPU tests)."""
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
    shells, shell2atom = [], []
    for a, ((Z, pos), (el, _)) in enumerate(zip(atoms, d["atoms_angstrom"])):
        for s in d["basis"][el]:
            shells.append(la.Shell(s["l"], list(zip(s["exponents"], s["coefficients"])), tuple(pos),
                                   pure=s["l"] > 1))
            shell2atom.append(a)
    water.shell2atom = shell2atom
    water.forces = d["forces"]
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
    b = la.Basis(ctx, la.BasisArrays.from_shells(shells))
    S, T, V = la.one_body(b, [(float(Z), pos) for Z, pos in atoms])  # pure d transformed on the GPU
    H = T + V
    fb = la.FockBuilder(b)

    def g(D):
        G, _, _ = fb.build(D, 1e-15)
        return G

    e, D, W = rhf(S, H, 5, g, enuc, diis=8, orbitals=True)
    print(f"RHF/aug-cc-pVDZ rotated water on the GPU Fock build: {e:.12f} (reference {eref})", flush=True)
    assert abs(e - eref) < tol, e
    # two-electron part of the energy gradient (the reference prints it as
    # "2-body forces"): tr(G1_x D) with G1 from the GPU derivative ERIs
    from libint_amd import deriv
    G1 = deriv.fock_deriv1(ctx, la.BasisArrays.from_shells(shells), water.shell2atom, len(atoms), D, 1e-15)
    F2 = np.array([np.sum(G1[x] * D) for x in range(3 * len(atoms))])
    ref = water.forces["F2"]
    print("  2-body forces", F2, flush=True)
    assert np.abs(F2 - ref["values"]).max() < ref["tolerance"], F2 - ref["values"]
    # 1-body, Pulay and total forces with the one-body derivatives from the GPU
    S1, T1, V1 = deriv.onebody_deriv1(ctx, la.BasisArrays.from_shells(shells), water.shell2atom, len(atoms),
                                      [(float(Z), pos) for Z, pos in atoms])
    dev = forces_check(S1, T1, V1, F2, D, W, atoms)
    print("  max deviation from the reference's 1-body, Pulay and total forces:", dev, flush=True)


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


def nuclear_repulsion_gradient(atoms):
    """d E_nuc / d x, as hartree-fock++.cc:672-701 forms it."""
    F = np.zeros((len(atoms), 3))
    for i in range(len(atoms)):
        for j in range(i):
            r = atoms[i][1] - atoms[j][1]
            f = -r * atoms[i][0] * atoms[j][0] / np.linalg.norm(r) ** 3
            F[i] += f
            F[j] -= f
    return F.reshape(-1)


def test_nuclear_repulsion_force():
    """Pins the geometry (units, atom order) at the reference's 1e-10."""
    atoms = water()[0]
    ref = water.forces["FN"]
    assert np.abs(nuclear_repulsion_gradient(atoms) - ref["values"]).max() < ref["tolerance"]


def test_two_body_force_oracle():
    """The reference's "2-body forces" golden with the oracle's derivative
    ERIs in deriv.fock_deriv1 (the Fock-derivative digest the GPU run uses),
    and by an independent route: Richardson-extrapolated central differences
    of tr(D G(D)) in the atom positions at fixed D."""
    import oracle
    import oracle.deriv1
    from libint_amd import deriv
    atoms, shells, enuc, eref, tol = water()
    S, H = one_body_pure(atoms, shells)
    arrays = la.BasisArrays.from_shells(shells)

    def g(D, arr=arrays):
        G, _, _, _ = oracle.fock_2e(arr, D, precision=1e-15)
        return G

    e, D = rhf(S, H, 5, g, enuc, diis=8, density=True)
    assert abs(e - eref) < tol
    ref = water.forces["F2"]
    G1 = deriv.fock_deriv1(None, arrays, water.shell2atom, len(atoms), D, 1e-15,
                           eri_deriv=lambda a, q, p, s: oracle.deriv1.eri_deriv1(a, q, p, s))
    F2 = np.array([np.sum(G1[x] * D) for x in range(3 * len(atoms))])
    assert np.abs(F2 - ref["values"]).max() < ref["tolerance"], F2 - ref["values"]

    def e2(a, x, h):
        shift = np.eye(3)[x] * h
        moved = [la.Shell(s.l, list(zip(s.alpha, s.coeff)), tuple(s.O + shift * (at == a)), pure=s.pure,
                          renorm=False) for s, at in zip(shells, water.shell2atom)]
        return float(np.sum(D * g(D, la.BasisArrays.from_shells(moved))))

    fd = []
    for a in range(len(atoms)):
        for x in range(3):
            h = 1e-3
            d1 = (e2(a, x, h) - e2(a, x, -h)) / (2 * h)
            d2 = (e2(a, x, h / 2) - e2(a, x, -h / 2)) / h
            fd.append((4 * d2 - d1) / 3)
    assert np.abs(np.array(fd) - ref["values"]).max() < 1e-7, np.array(fd) - ref["values"]


def _oracle_one_body_fn(aug, charges):
    """oracle one-body matrices over a Cartesian BasisArrays (the augmented basis)."""
    off = np.concatenate([[0], np.cumsum(aug.nprim)])
    sh = [(int(aug.l[s]), aug.alpha[off[s]:off[s + 1]], aug.coeff[off[s]:off[s + 1]], aug.origin[3 * s:3 * s + 3])
          for s in range(aug.nshell)]
    return onebody.one_body(sh, [(float(Z), np.asarray(p, float)) for Z, p in charges])


def forces_check(S1, T1, V1, F2, D, W, atoms):
    f = water.forces
    nx = 3 * len(atoms)
    F1 = np.array([2 * np.sum((T1[x] + V1[x]) * D) for x in range(nx)])
    FP = np.array([-2 * np.sum(S1[x] * W) for x in range(nx)])
    FN = nuclear_repulsion_gradient(atoms)
    F = F1 + FP + F2 + FN
    out = {}
    for name, v in (("F1", F1), ("FP", FP), ("F", F)):
        out[name] = np.abs(v - f[name]["values"]).max()
        assert out[name] < f[name]["tolerance"], (name, v - f[name]["values"])
    return out


def test_all_forces_oracle():
    """The reference's 1-body, Pulay and total Hartree-Fock force goldens
    (hartree-fock++-validate.py:74-101) from the oracle's one-body and ERI
    derivatives through deriv.onebody_deriv1 / deriv.fock_deriv1."""
    import oracle
    import oracle.deriv1
    from libint_amd import deriv
    atoms, shells, enuc, eref, tol = water()
    S, H = one_body_pure(atoms, shells)
    arrays = la.BasisArrays.from_shells(shells)

    def g(D):
        return oracle.fock_2e(arrays, D, precision=1e-15)[0]

    e, D, W = rhf(S, H, 5, g, enuc, diis=8, orbitals=True)
    assert abs(e - eref) < tol
    charges = [(float(Z), pos) for Z, pos in atoms]
    S1, T1, V1 = deriv.onebody_deriv1(None, arrays, water.shell2atom, len(atoms), charges,
                                      one_body_fn=_oracle_one_body_fn)
    G1 = deriv.fock_deriv1(None, arrays, water.shell2atom, len(atoms), D, 1e-15,
                           eri_deriv=lambda a, q, p, s: oracle.deriv1.eri_deriv1(a, q, p, s))
    F2 = np.array([np.sum(G1[x] * D) for x in range(3 * len(atoms))])
    forces_check(S1, T1, V1, F2, D, W, atoms)
