"""Basis sets from Gaussian94 files and molecules from .xyz files, with the
semantics of libint2::BasisSet (include/libint2/basis.h.in) and
libint2::read_dotxyz (include/libint2/atom.h).

* BasisSet(name, atoms): the file <data path>/<canonical name>.g94 (or its
  components: aug-cc-pvXz* = cc-pvXz* + augmentation-cc-pvXz*,
  basis.h.in:388-400), each atom receiving its element's shells in file
  order, component after component (:134-200).  Names are lower-cased and
  '/' becomes 'I' (:345-364).  d shells are pure except in the Pople sets
  that Gaussian runs with Cartesian d (3-21*, 4-31G, 6-31*G, 6-31*G*,
  :366-385); f and higher are always pure (:568).
* The data path is $LIBINT_DATA_PATH/basis, else $LIBINT_DATA_PATH itself
  (:412-449); an explicit data_path argument takes its place.
* read_g94: '!' comments, EMSL ('****' first) and MolSSI layouts, Fortran
  'D' exponents, SP shells split into an s and a p shell (:460-622).
* read_dotxyz: natoms, a comment line, then symbol x y z in Angstrom,
  converted with bohr_to_angstrom = 0.529177210903 (CODATA 2018, atom.h:
  50-55, 80-140, 209-216).
Shells are la.Shell objects (coefficients renormalised as libint does).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import BasisArrays, Shell

_SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
            "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
            "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn").split()
Z_OF = {s.lower(): i + 1 for i, s in enumerate(_SYMBOLS)}
# angular momentum labels (Shell::am_symbol_to_l); Gaussian writes J for
# l = 7 and shifts K, L, M ... by one (basis.h.in:566-575)
_AM = "spdfghiklmnoqrtuvwxyz"

BOHR_TO_ANGSTROM = 0.529177210903


@dataclass
class Atom:
    atomic_number: int
    x: float
    y: float
    z: float


def canonicalize_name(name: str) -> str:
    return "".join("I" if c == "/" else c for c in name.lower())


def gaussian_cartesian_d_convention(canonical_name: str) -> bool:
    n = canonical_name
    if n.startswith("3-21") or n.startswith("4-31g"):
        return True
    if n.startswith("6-31") and len(n) > 4 and n[4] != "1":
        g = n.find("g")
        if g < 0:
            return False
        if g + 1 == len(n):
            return True
        if n[g + 1] in "*s":
            return True
    return False


def decompose_name_into_components(canonical_name: str) -> list[str]:
    if canonical_name.startswith("aug-cc-pv") and "cabs" not in canonical_name:
        base = canonical_name[4:]
        return [base, "augmentation-" + base]
    return [canonical_name]


def data_path(path: str | None = None) -> str:
    path = path if path is not None else os.environ.get("LIBINT_DATA_PATH")
    if not path:
        raise OSError("no basis-set data path: set LIBINT_DATA_PATH or pass data_path")
    for p in (os.path.join(path, "basis"), path):
        if os.path.isdir(p):
            return p
    raise OSError(f'BasisSet data path "{path}{{/basis}}" is not valid')


def _l_of(label: str) -> int:
    c = label[0].lower()
    if c == "j":
        return 7
    l = _AM.index(c)
    return l + 1 if l >= 7 else l


def read_g94(path: str, force_cartesian_d: bool = False, throw_if_missing: bool = True) -> dict:
    """{Z: [(l, pure, exponents, coefficients), ...]} in file order."""
    if not os.path.exists(path):
        if throw_if_missing:
            raise OSError(f'could not open "{path}"')
        return {}
    lines = [ln.rstrip("\n") for ln in open(path)]
    lines = [ln for ln in lines if ln.strip() and not ln.lstrip().startswith("!")]
    out: dict = {}
    i = 0
    first = True
    expect_element = True
    Z = None
    while i < len(lines):
        ln = lines[i].strip()
        i += 1
        if ln == "****":
            if first:
                continue
            expect_element = True
            continue
        if expect_element:
            expect_element = False
            first = False
            sym = ln.split()[0].lower()
            if sym not in Z_OF:
                raise ValueError(f'in file {path} found G94 basis set for element symbol "{sym}", '
                                 f"not found in the periodic table")
            Z = Z_OF[sym]
            out.setdefault(Z, [])
            continue
        tok = ln.split()
        label, nprim = tok[0], int(tok[1])
        rows = []
        for _ in range(nprim):
            rows.append([float(x) for x in lines[i].replace("D", "E").replace("d", "e").split()])
            i += 1
        ex = [r[0] for r in rows]
        if label.lower() == "sp":
            out[Z].append((0, False, ex, [r[1] for r in rows]))
            out[Z].append((1, False, ex, [r[2] for r in rows]))
        else:
            l = _l_of(label)
            pure = l > 2 if force_cartesian_d else l > 1
            out[Z].append((l, pure, ex, [r[1] for r in rows]))
    return out


def read_dotxyz(text: str, bohr_to_angstrom: float = BOHR_TO_ANGSTROM) -> list[Atom]:
    """Atoms of an .xyz text (positions converted to bohr)."""
    ln = [l for l in text.splitlines()]
    n = int(ln[0].split()[0])
    atoms = []
    for k in range(n):
        t = ln[2 + k].split()
        sym = t[0].lower()
        if sym not in Z_OF:
            raise ValueError(f'element symbol "{t[0]}" is not recognized')
        x, y, z = (float(v) / bohr_to_angstrom for v in t[1:4])
        atoms.append(Atom(Z_OF[sym], x, y, z))
    return atoms


class BasisSet(list):
    """The shells of `name` on `atoms`, as libint2::BasisSet(name, atoms)."""

    def __init__(self, name: str, atoms, throw_if_no_match: bool = False, data_path_override: str | None = None):
        super().__init__()
        self.name = name
        path = data_path(data_path_override)
        canon = canonicalize_name(name)
        cart_d = gaussian_cartesian_d_convention(canon)
        comps = decompose_name_into_components(canon)
        libs = [read_g94(os.path.join(path, c + ".g94"), cart_d, throw_if_no_match) for c in comps]
        self._atom_of = []
        for a, atom in enumerate(atoms):
            for c, lib in zip(comps, libs):
                shells = lib.get(atom.atomic_number, [])
                if not shells and throw_if_no_match:
                    raise LookupError(f"did not find the basis for this Z in {os.path.join(path, c + '.g94')}")
                for l, pure, ex, co in shells:
                    self.append(Shell(l, list(zip(ex, co)), (atom.x, atom.y, atom.z), pure=pure))
                    self._atom_of.append(a)

    def nbf(self) -> int:
        return int(sum(s.size() for s in self))

    def max_nprim(self) -> int:
        return max((s.nprim() for s in self), default=0)

    def max_l(self) -> int:
        return max((s.l for s in self), default=-1)

    def shell2bf(self) -> list[int]:
        out, n = [], 0
        for s in self:
            out.append(n)
            n += s.size()
        return out

    def shell2atom(self) -> list[int]:
        return list(self._atom_of)

    def arrays(self) -> BasisArrays:
        return BasisArrays.from_shells(self)
