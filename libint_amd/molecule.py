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


def screened_quartets(p1: np.ndarray, p2: np.ndarray, K: np.ndarray, thresh: float) -> np.ndarray:
    """Permutationally-unique quartets over a shell-pair list (pair index
    a >= b in the (s1, s2) loop order) whose Schwarz bound K_a K_b is at least
    thresh, as an (n, 4) int32 array."""
    order = np.lexsort((p2, p1))  # loop order s1, then s2
    p1, p2, K = p1[order], p2[order], K[order]
    chunks = []
    for a in range(K.size):
        if K[a] == 0:
            continue
        b = np.nonzero(K[:a + 1] * K[a] >= thresh)[0]
        if b.size:
            chunks.append(np.stack([np.full(b.size, p1[a]), np.full(b.size, p2[a]), p1[b], p2[b]], 1))
    if not chunks:
        return np.zeros((0, 4), np.int32)
    return np.concatenate(chunks).astype(np.int32)


def headline_quartets(basis, thresh: float = 1e-12):
    """The (dd|dd) work list of the 60-atom batched workload: every
    permutationally-unique quartet of d shells whose Schwarz bound
    K12 K34 >= thresh, with K(s1, s2) = sqrt(max|(s1 s2|s1 s2)|) computed on
    the GPU without primitive screening (compute_schwarz_ints,
    tests/hartree-fock/hartree-fock++.cc:1231-1298).  `basis` is a device
    Basis.  Returns (quartets, number of unique (dd|dd) quartets before
    screening)."""
    from . import schwarz
    dsh = class_shells(basis.arrays, 2)
    i, j = np.tril_indices(dsh.size)
    p1, p2 = dsh[i], dsh[j]
    K = schwarz(basis, np.stack([p1, p2], 1), precision=0.0)
    return screened_quartets(p1, p2, K, thresh), K.size * (K.size + 1) // 2


def stratified_quartets(arrays: BasisArrays, pairs: np.ndarray, K: np.ndarray, per_class: int,
                        rng: np.random.Generator, thresh: float = 1e-12, shuffle_order: bool = True,
                        per_class_of=None) -> np.ndarray:
    """Seeded sample of quartets (ab|cd) whose Schwarz bound K_ab K_cd is at
    least thresh, stratified by class: for every combination of a bra pair
    type and a ket pair type (pair type = the two l of the pair), up to
    per_class quartets (per_class_of(la, lb, lc, ld) overrides it) drawn
    uniformly among the significant ones.  With shuffle_order each quartet
    comes in a random caller order (bra/ket and within-pair swaps), so that
    the reordering and un-permuting are exercised too.  pairs: (n, 2) shell
    pairs, K their Schwarz factors."""
    pairs = np.asarray(pairs)
    lp = np.sort(arrays.l[pairs], axis=1)[:, ::-1]
    types = sorted({tuple(t) for t in lp.tolist()})
    members = {t: np.nonzero((lp[:, 0] == t[0]) & (lp[:, 1] == t[1]))[0] for t in types}
    out = []
    for ti, t1 in enumerate(types):
        for t2 in types[ti:]:
            want = per_class if per_class_of is None else per_class_of(t1[0], t1[1], t2[0], t2[1])
            if want <= 0:
                continue
            A, Bm = members[t1], members[t2]
            got = []
            n = 0
            for _ in range(50):
                a = A[rng.integers(0, A.size, 4 * want)]
                b = Bm[rng.integers(0, Bm.size, 4 * want)]
                ok = K[a] * K[b] >= thresh
                got.append(np.stack([a[ok], b[ok]], 1))
                n += int(ok.sum())
                if n >= want:
                    break
            ab = np.concatenate(got)[:want]
            if ab.size:
                out.append(np.concatenate([pairs[ab[:, 0]], pairs[ab[:, 1]]], 1))
    q = np.concatenate(out).astype(np.int32) if out else np.zeros((0, 4), np.int32)
    if shuffle_order and len(q):
        sb = rng.integers(0, 2, len(q)).astype(bool)
        sk = rng.integers(0, 2, len(q)).astype(bool)
        q[sb, 0], q[sb, 1] = q[sb, 1].copy(), q[sb, 0].copy()
        q[sk, 2], q[sk, 3] = q[sk, 3].copy(), q[sk, 2].copy()
        bk = rng.integers(0, 2, len(q)).astype(bool)
        q[bk] = q[bk][:, [2, 3, 0, 1]]
    return q


def sampled_row_shells(mol: Molecule) -> list[int]:
    """Shells whose rows of a large J/K build are kept for checking: the
    first s, d and f shells of the atom nearest the centre of the molecule,
    then the first f shell of the atom farthest from it (the fewest
    quartets touch it, so its exact rows are the cheapest to compute)."""
    B = mol.basis
    c = mol.coords.mean(0)
    d2 = ((mol.coords - c) ** 2).sum(1)
    near, far = int(np.argmin(d2)), int(np.argmax(d2))
    pick = []
    for l in (0, 2, 3):
        pick += [s for s in range(B.nshell) if mol.atom_of_shell[s] == near and B.l[s] == l][:1]
    pick += [s for s in range(B.nshell) if mol.atom_of_shell[s] == far and B.l[s] == 3][:1]
    return pick
