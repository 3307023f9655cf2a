"""First geometric derivatives of the Coulomb ERIs (deriv_order = 1).

libint's Engine(Operator::coulomb, max_nprim, max_l, deriv_order=1) returns,
per shell quartet, 12 shell sets: d/dA_x, d/dA_y, d/dA_z, d/dB_x, ..., d/dD_z
for the centres of the shells in caller order (the order of
CartesianDerivIterator<4>, include/libint2/util/intpart_iter.h:89-103,
161-182), each row-major in caller order like the integrals themselves.

eri_deriv1 runs the device derivative plan (l2g_eri_deriv1_plan_*): the ERIs
with one shell raised (l+1, coefficients times 2 alpha) and lowered (l-1)
in one ERI plan on the same kernels, and a combination kernel that forms
d/dA_i = raised - a_i lowered per Cartesian component and applies the
solid-harmonic rows of the pure centres.  fock_deriv1 digests the
derivatives into the derivative Fock matrices as the reference's
Hartree-Fock driver does.
"""
from __future__ import annotations

import numpy as np

from . import SCREEN_ORIGINAL, Basis, BasisArrays, Deriv1Plan


def eri_deriv1(ctx, arrays: BasisArrays, quartets, precision=np.finfo(float).eps, screening=SCREEN_ORIGINAL):
    """First derivatives of (s0 s1|s2 s3) for each quartet (caller order).

    Returns (values, offsets, nonzero): values[d] holds the 12 derivative
    shell sets d = 3 * centre + axis back to back, quartet q's set at
    values[d][offsets[q]: offsets[q] + size_q]; nonzero[q] is 0 when every
    primitive of the raised and lowered quartets was screened out."""
    p = Deriv1Plan(Basis(ctx, arrays), quartets, precision, screening)
    offs = p.offsets()
    p.execute()
    vals, nz = p.download()
    return vals, offs, nz


def fock_deriv1(ctx, arrays: BasisArrays, shell2atom, natoms, D, precision=np.finfo(float).eps,
                screening=SCREEN_ORIGINAL, eri_deriv=None):
    """First derivatives of the two-electron Fock matrix, as the reference's
    compute_2body_fock_deriv<1> forms them (tests/hartree-fock/
    hartree-fock++.cc:1776-1990): over the permutationally unique quartets
    (s1 >= s2, s3 <= s1, s4 <= (s1 == s3 ? s2 : s3)) each of the 12 derivative
    shell sets, weighted by the quartet's degeneracy, is added to the matrix
    of the coordinate of its centre's atom with the six-way rule
    g(1,2) += D(3,4) v, g(3,4) += D(1,2) v, g(1,3), g(2,4), g(1,4), g(2,3)
    -= D/4 v, and each matrix is symmetrised at the end, G = (g + g^T)/2.

    shell2atom[s] is the atom of shell s.  Returns G of shape
    (3 * natoms, nbf, nbf), coordinate 3 * atom + xyz; sum(G[x] * D) is the
    two-electron part of the energy gradient, d tr(D G(D)) / dx at fixed D.
    eri_deriv(arrays, quartets, precision, screening) -> (values, offsets,
    nonzero) replaces the device derivative plan (a checker passes its own)."""
    ns = arrays.nshell
    sizes = arrays.sizes().astype(np.int64)
    bf = np.concatenate([[0], np.cumsum(sizes)])
    nbf = int(bf[-1])
    D = np.asarray(D, float)
    quartets = [(a, b, c, d) for a in range(ns) for b in range(a + 1) for c in range(a + 1)
                for d in range((b if a == c else c) + 1)]
    q = np.array(quartets, np.int32).reshape(-1, 4)
    if eri_deriv is None:
        vals, off, nz = eri_deriv1(ctx, arrays, q, precision, screening)
    else:
        vals, off, nz = eri_deriv(arrays, q, precision, screening)
    atom = np.asarray(shell2atom, np.int64)
    G = np.zeros((3 * natoms, nbf, nbf))
    if len(q) == 0:
        return G
    s1, s2, s3, s4 = (q[:, k].astype(np.int64) for k in range(4))
    deg = np.where(s1 == s2, 1.0, 2.0) * np.where(s3 == s4, 1.0, 2.0) * \
        np.where((s1 == s3) & (s2 == s4), 1.0, 2.0)
    dims = sizes[q]
    key = ((dims[:, 0] * 16 + dims[:, 1]) * 16 + dims[:, 2]) * 16 + dims[:, 3]
    keep = nz != 0
    # quartets of equal block shape in one batch: the six scatter terms as
    # batched contractions, added into G with np.add.at
    for kv in np.unique(key[keep]):
        idx = np.nonzero((key == kv) & keep)[0]
        n = [int(x) for x in dims[idx[0]]]
        m = len(idx)
        blk = vals[:, off[idx][:, None] + np.arange(int(np.prod(n)))]  # (12, m, size)
        V = blk.transpose(1, 0, 2).reshape([m, 12] + n) * deg[idx][:, None, None, None, None, None]
        r = [bf[sk[idx]][:, None] + np.arange(nk) for sk, nk in zip((s1, s2, s3, s4), n)]
        coord = 3 * atom[q[idx][:, np.arange(12) // 3]] + np.arange(12) % 3  # (m, 12)

        def Dblk(i, j):
            return D[r[i][:, :, None], r[j][:, None, :]]  # (m, n_i, n_j)

        def add(t, i, j, scale):
            np.add.at(G, (coord[:, :, None, None], r[i][:, None, :, None], r[j][:, None, None, :]), scale * t)

        add(np.einsum("mxabcd,mcd->mxab", V, Dblk(2, 3)), 0, 1, 1.0)
        add(np.einsum("mxabcd,mab->mxcd", V, Dblk(0, 1)), 2, 3, 1.0)
        add(np.einsum("mxabcd,mbd->mxac", V, Dblk(1, 3)), 0, 2, -0.25)
        add(np.einsum("mxabcd,mac->mxbd", V, Dblk(0, 2)), 1, 3, -0.25)
        add(np.einsum("mxabcd,mbc->mxad", V, Dblk(1, 2)), 0, 3, -0.25)
        add(np.einsum("mxabcd,mad->mxbc", V, Dblk(0, 3)), 1, 2, -0.25)
    return 0.5 * (G + G.transpose(0, 2, 1))


def _cart_exps(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def _cart_index(x, y, z):
    l = x + y + z
    return (l - x) * (l - x + 1) // 2 + (l - x - y)


def onebody_deriv1(ctx, arrays: BasisArrays, shell2atom, natoms, charges, one_body_fn=None):
    """First derivatives of the overlap, kinetic and point-charge attraction
    matrices with respect to the atom coordinates, as the reference's
    compute_1body_ints_deriv<overlap / kinetic / nuclear>(1, obs, atoms)
    forms them (tests/hartree-fock/hartree-fock++.cc:1028-1227): S1, T1, V1 of
    shape (3 * natoms, nbf, nbf), coordinate 3 * atom + xyz.  charges[c] =
    (Z, position) sits on atom c (the nuclei, in atom order).

    The bra derivative d/dA_i <a|O|b> = <a+1_i|O|b> (coefficients 2 alpha c)
    - n_i <a-1_i|O|b> comes from the one-body matrices of an augmented basis
    (Cartesian copies, raised, lowered shells) on the GPU; the ket derivative
    is its transpose (S, T, V are symmetric); the derivative of each charge's
    own term with respect to the charge's position is minus the sum of the
    bra and ket derivatives of that term (translational invariance).  Pure
    shells are transformed at the end.  one_body_fn(arrays, charges) ->
    (S, T, V) replaces la.one_body (a checker passes its own)."""
    from . import Basis, one_body, solid_harmonics_coeff
    n = arrays.nshell
    L = arrays.l.astype(int)
    if len(charges) != natoms:
        raise ValueError("onebody_deriv1 takes the nuclei as the charges, one per atom in atom order")
    if L.max(initial=0) > 2:
        raise ValueError("one-body derivatives are provided up to l = 2 (raised shells up to 3)")
    off = np.concatenate([[0], np.cumsum(arrays.nprim)])
    ls, pure, nprim, al, co, org = [], [], [], [], [], []
    for kind in range(3):
        for s in range(n):
            a = arrays.alpha[off[s]:off[s + 1]]
            c = arrays.coeff[off[s]:off[s + 1]]
            l = int(L[s])
            if kind == 1:
                l, c = l + 1, 2.0 * a * c
            elif kind == 2:
                l = max(l - 1, 0)
            ls.append(l)
            pure.append(0)
            nprim.append(len(a))
            al.append(a)
            co.append(c)
            org.append(arrays.origin[3 * s:3 * s + 3])
    aug = BasisArrays(ls, pure, nprim, np.concatenate(al), np.concatenate(co), np.concatenate(org))
    if one_body_fn is None:
        b = Basis(ctx, aug)

        def one_body_fn(_arrays, ch):
            return one_body(b, ch)
    nca = [(l + 1) * (l + 2) // 2 for l in ls]
    aoff = np.concatenate([[0], np.cumsum(nca)])
    ncart = int(aoff[n])  # the Cartesian copies come first

    def bra_derivs(M):
        """d/dA_i <a|.|b> over Cartesian a (rows) and b (columns), i = x, y, z."""
        out = np.zeros((3, ncart, ncart))
        for s in range(n):
            l = int(L[s])
            for ci, e in enumerate(_cart_exps(l)):
                row = aoff[s] + ci
                for i in range(3):
                    up = list(e)
                    up[i] += 1
                    ru = aoff[n + s] + _cart_index(*up)
                    v = M[ru, :ncart].copy()
                    if e[i] > 0:
                        dn = list(e)
                        dn[i] -= 1
                        v -= e[i] * M[aoff[2 * n + s] + _cart_index(*dn), :ncart]
                    out[i, row] = v
        return out

    atom_of_fn = np.concatenate([[shell2atom[s]] * nca[s] for s in range(n)]) if n else np.zeros(0, int)
    # Cartesian -> the shells' own functions
    blocks = []
    for s in range(n):
        l = int(L[s])
        if arrays.pure[s] and l > 0:
            blocks.append(np.array([[solid_harmonics_coeff(l, m, *c) for c in _cart_exps(l)]
                                    for m in range(-l, l + 1)]).T)
        else:
            blocks.append(np.eye((l + 1) * (l + 2) // 2))
    X = np.zeros((ncart, sum(bk.shape[1] for bk in blocks)))
    r = c = 0
    for bk in blocks:
        X[r:r + bk.shape[0], c:c + bk.shape[1]] = bk
        r += bk.shape[0]
        c += bk.shape[1]

    def assemble(Dbra, extra=None):
        """Per atom coordinate: bra part on rows of that atom plus its
        transpose, plus the operator-centre term; then the pure transform."""
        out = np.zeros((3 * natoms, X.shape[1], X.shape[1]))
        for a in range(natoms):
            mask = (atom_of_fn == a).astype(float)[:, None]
            for i in range(3):
                M = Dbra[i] * mask
                M = M + M.T
                if extra is not None:
                    M = M + extra[3 * a + i]
                out[3 * a + i] = X.T @ M @ X
        return out

    S, T, _ = one_body_fn(aug, [])
    S1 = assemble(bra_derivs(S))
    T1 = assemble(bra_derivs(T))
    V1 = np.zeros_like(S1)
    for cidx, (Z, pos) in enumerate(charges):
        _, _, Vc = one_body_fn(aug, [(Z, pos)])
        Dc = bra_derivs(Vc)
        full = Dc + Dc.transpose(0, 2, 1)  # d/dA + d/dB of the charge's term, all rows
        extra = np.zeros((3 * natoms, ncart, ncart))
        for i in range(3):
            extra[3 * cidx + i] = -full[i]
        V1 += assemble(Dc, extra)
    return S1, T1, V1
