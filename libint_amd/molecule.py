"""Seeded synthetic molecules and quartet lists for the five workloads.

Atoms are placed uniformly at random in a cube (rejection sampling with a
minimum separation), each atom carries a fixed recipe of shells with
log-uniform exponents and U[0.1, 1] contraction coefficients that are then
renormalised (libint2::Shell::renorm, include/libint2/shell.h:958-999).
d and f shells are pure (real solid harmonics, Standard order), s and p are
Cartesian.  Positions are drawn in Angstrom and converted to bohr with
CODATA 2018 (include/libint2/atom.h:52-55).

Unique quartets follow the permutationally-unique loop of libint's Fock
builders, s1 >= s2, s3 <= s1, s4 <= (s1 == s3 ? s2 : s3)
(tests/hartree-fock/hartree-fock.cc:677-692).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import BasisArrays, shell_renorm

ANGSTROM_PER_BOHR = 0.529177210903

# (natoms, cube edge in Angstrom, per-atom recipe [(l, K spec)], exponent range)
#   K spec: an int (fixed) or ("U", kmax) for uniform {1..kmax}
WORKLOADS = {
    1: dict(natoms=10, box=6.0, recipe=[(0, 6), (0, 1), (0, 1), (1, 3), (1, 1), (2, 1)], expo=(0.1, 5e3)),
    2: dict(natoms=60, box=14.0, recipe=[(0, ("U", 6))] * 4 + [(1, ("U", 6))] * 3 + [(2, ("U", 6))] * 2,
            expo=(0.05, 1e4)),
    3: dict(natoms=120, box=20.0,
            recipe=[(0, ("U", 8))] * 4 + [(1, ("U", 8))] * 3 + [(2, ("U", 8))] * 2 + [(3, ("U", 8))],
            expo=(0.03, 2e4)),
    4: dict(natoms=200, box=26.0,
            recipe=[(0, ("U", 8))] * 4 + [(1, ("U", 8))] * 3 + [(2, ("U", 8))] * 2 + [(3, ("U", 8))],
            expo=(0.03, 2e4)),
}


@dataclass
class Molecule:
    coords: np.ndarray       # (natoms, 3) bohr
    basis: BasisArrays
    atom_of_shell: np.ndarray
    workload: int

    @property
    def nbf(self) -> int:
        return int(self.basis.sizes().sum())


def place_atoms(rng: np.random.Generator, natoms: int, box: float, min_sep: float = 0.9) -> np.ndarray:
    pts = []
    tries = 0
    while len(pts) < natoms:
        p = rng.uniform(0.0, box, 3)
        tries += 1
        if tries > 1000000:
            raise RuntimeError("could not place atoms with the requested separation")
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
    return np.array(pts) / ANGSTROM_PER_BOHR


def make_molecule(workload: int, natoms: int | None = None, seed: int | None = None,
                  max_l: int | None = None, same_density: bool = False) -> Molecule:
    """Molecule of a workload recipe.  seed defaults to 1000 + workload.
    max_l drops shells above that angular momentum (for reduced tests).
    same_density scales the cube edge with natoms^(1/3) so that a smaller
    molecule keeps the recipe's atom density (else the edge is the recipe's)."""
    spec = WORKLOADS[workload]
    n = spec["natoms"] if natoms is None else natoms
    rng = np.random.Generator(np.random.PCG64(1000 + workload if seed is None else seed))
    coords = place_atoms(rng, n, box_edge(workload, n) if same_density else spec["box"])
    lo, hi = spec["expo"]
    l_, pure_, np_, al_, co_, org_, atom_ = [], [], [], [], [], [], []
    for a in range(n):
        for (l, kspec) in spec["recipe"]:
            K = kspec if isinstance(kspec, int) else int(rng.integers(1, kspec[1] + 1))
            alpha = np.exp(rng.uniform(np.log(lo), np.log(hi), K))
            coeff = rng.uniform(0.1, 1.0, K)
            if max_l is not None and l > max_l:
                continue
            c, _ = shell_renorm(l, alpha, coeff)
            l_.append(l)
            pure_.append(1 if l >= 2 else 0)
            np_.append(K)
            al_.append(alpha)
            co_.append(c)
            org_.append(coords[a])
            atom_.append(a)
    b = BasisArrays(l_, pure_, np_, np.concatenate(al_), np.concatenate(co_), np.array(org_))
    return Molecule(coords, b, np.array(atom_), workload)


def box_edge(workload: int, natoms: int) -> float:
    """Cube edge (Angstrom) holding natoms at the recipe's atom density."""
    spec = WORKLOADS[workload]
    return spec["box"] * (natoms / spec["natoms"]) ** (1.0 / 3.0)


def unique_quartets(nshell: int, shells: np.ndarray | None = None) -> np.ndarray:
    """All permutationally-unique quartets over `shells` (default: all),
    as an (n, 4) int32 array of shell indices."""
    idx = np.arange(nshell) if shells is None else np.asarray(shells)
    # pairs (s1 >= s2)
    i, j = np.tril_indices(idx.size)
    p1, p2 = idx[i], idx[j]
    npair = p1.size
    # quartets = pairs of pairs with pair index (ij) >= (kl) in the loop order
    a, b = np.tril_indices(npair)
    q = np.stack([p1[a], p2[a], p1[b], p2[b]], axis=1).astype(np.int32)
    return q


def quartet_count(nshell: int) -> int:
    npair = nshell * (nshell + 1) // 2
    return npair * (npair + 1) // 2


def class_shells(basis: BasisArrays, l: int) -> np.ndarray:
    return np.nonzero(basis.l == l)[0]
