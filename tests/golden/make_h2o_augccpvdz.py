"""Writes tests/golden/h2o_augccpvdz.json: the aug-cc-pVDZ shells of H and O
as libint's BasisSet("aug-cc-pVDZ") assembles them from lib/basis/cc-pvdz.g94
and lib/basis/augmentation-cc-pvdz.g94 (include/libint2/basis.h.in:134-200,
388-400; d shells pure, :568) and the water geometry of
tests/hartree-fock/h2o_rotated.xyz, for the end-to-end Hartree-Fock check
against the reference's -76.003354058439 and its force components
(tests/hartree-fock/hartree-fock++-validate.py:47-48, 74-103: 1-body, Pulay,
2-body, nuclear-repulsion and total forces, read from that file's text).
Data only; run here
with the reference checked out:
    python tests/golden/make_h2o_augccpvdz.py /root/reference
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
L = {"S": 0, "P": 1, "D": 2, "F": 3}


def read_g94(path, elements):
    out = {}
    lines = [l.strip() for l in open(path) if l.strip() and not l.startswith("!")]
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        if tok[0] in elements and len(tok) == 2:
            el = tok[0]
            shells = []
            i += 1
            while lines[i] != "****":
                kind, n = lines[i].split()[:2]
                n = int(n)
                prims = [lines[i + 1 + k].replace("D", "E").split() for k in range(n)]
                shells.append({"l": L[kind], "exponents": [float(p[0]) for p in prims],
                               "coefficients": [float(p[1]) for p in prims]})
                i += 1 + n
            out[el] = shells
        i += 1
    return out


# aug-cc-pVDZ = cc-pVDZ + augmentation-cc-pVDZ, shells in that order per atom
# (decompose_name_into_components, basis.h.in:388-400, and :159-170)
base = read_g94(os.path.join(ref, "lib", "basis", "cc-pvdz.g94"), {"H", "O"})
aug = read_g94(os.path.join(ref, "lib", "basis", "augmentation-cc-pvdz.g94"), {"H", "O"})
basis = {el: base[el] + aug[el] for el in ("H", "O")}
xyz = open(os.path.join(ref, "tests", "hartree-fock", "h2o_rotated.xyz")).read().split("\n")
natom = int(xyz[0])
atoms = [(l.split()[0], [float(x) for x in l.split()[1:4]]) for l in xyz[2:2 + natom]]
val = open(os.path.join(ref, "tests", "hartree-fock", "hartree-fock++-validate.py")).read()


def vector(name):
    return [float(x) for x in re.search(name + r"\s*=\s*\[([^\]]*)\]", val).group(1).split(",")]


def number(name):
    return float(re.search(name + r"\s*=\s*([-+0-9.e]+)", val).group(1))


forces = {k: {"values": vector(k + "ref"), "tolerance": number(k + "tol")} for k in ("F1", "FP", "F2", "FN", "F")}
assert vector("eref") == [-76.003354058439] and number("etol") == 5e-12
json.dump({"basis": basis, "atoms_angstrom": atoms, "energy": -76.003354058439, "energy_tolerance": 5e-12,
           "forces": forces},
          open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "h2o_augccpvdz.json"), "w"), indent=1)
