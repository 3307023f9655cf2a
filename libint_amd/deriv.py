"""First geometric derivatives of the Coulomb ERIs (deriv_order = 1).

libint's Engine(Operator::coulomb, max_nprim, max_l, deriv_order=1) returns,
per shell quartet, 12 shell sets: d/dA_x, d/dA_y, d/dA_z, d/dB_x, ..., d/dD_z
for the centres of the shells in caller order (the order of
CartesianDerivIterator<4>, include/libint2/util/intpart_iter.h:89-103,
161-182), each row-major in caller order like the integrals themselves.

Differentiating a Cartesian Gaussian with respect to its centre,

    d/dA_i  x^a_i e^{-alpha r^2} = 2 alpha x^(a_i+1) e^{...} - a_i x^(a_i-1) e^{...},

turns each derivative shell set into ERIs of shells with l raised by one (the
contraction coefficients multiplied by 2 alpha, exponents unchanged) and
lowered by one.  Those ERIs are evaluated on the GPU in one plan (the same
kernels, parity and precision as the integrals themselves), and combined per
Cartesian component; pure shells are transformed to solid harmonics after
the combination, which commutes with it.  The combination itself is a few
multiply-adds per derivative integral on the host.
"""
from __future__ import annotations

import numpy as np

from . import (SCREEN_ORIGINAL, Basis, BasisArrays, compute_quartets, solid_harmonics_coeff)


def _cart(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def _cart_index(x, y, z):
    l = x + y + z
    return (l - x) * (l - x + 1) // 2 + (l - x - y)


def _sph_matrix(l):
    """Rows m = -l..l (libint's standard ordering), columns Cartesian."""
    return np.array([[solid_harmonics_coeff(l, m, *c) for c in _cart(l)] for m in range(-l, l + 1)])


def _augmented(arrays: BasisArrays):
    """Cartesian copies of the shells, then the raised (l+1, coefficients
    2 alpha c) and the lowered (l-1, same coefficients; an unused s copy for
    l = 0) shells."""
    n = arrays.nshell
    off = np.concatenate([[0], np.cumsum(arrays.nprim)])
    ls, pure, nprim, al, co, org = [], [], [], [], [], []
    for kind in range(3):
        for s in range(n):
            a = arrays.alpha[off[s]:off[s + 1]]
            c = arrays.coeff[off[s]:off[s + 1]]
            l = int(arrays.l[s])
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
    return BasisArrays(ls, pure, nprim, np.concatenate(al), np.concatenate(co), np.concatenate(org))


def eri_deriv1(ctx, arrays: BasisArrays, quartets, precision=np.finfo(float).eps, screening=SCREEN_ORIGINAL):
    """First derivatives of (s0 s1|s2 s3) for each quartet (caller order).

    Returns (values, offsets, nonzero): values[d] holds the 12 derivative
    shell sets d = 3 * centre + axis back to back, quartet q's set at
    values[d][offsets[q]: offsets[q] + size_q]; nonzero[q] is 0 when every
    primitive of the raised and lowered quartets was screened out."""
    q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
    n = arrays.nshell
    aug = _augmented(arrays)
    b = Basis(ctx, aug)
    L = arrays.l
    # raised / lowered quartets per centre
    big = []
    for k in range(4):
        r = q.copy()
        r[:, k] += n
        big.append(r)
    for k in range(4):
        r = q.copy()
        r[:, k] += 2 * n
        big.append(r)
    allq = np.concatenate(big) if len(q) else np.zeros((0, 4), np.int32)
    vals, offs, nz = compute_quartets(b, allq, precision, screening)
    nq = len(q)
    size = np.ones(nq, np.int64)
    sizes = arrays.sizes()
    for k in range(4):
        size *= sizes[q[:, k]]
    out_off = np.concatenate([[0], np.cumsum(size)[:-1]]) if nq else np.zeros(0, np.int64)
    total = int(size.sum())
    out = np.zeros((12, max(total, 1)))
    if nq == 0:
        return out[:, :0], out_off, np.zeros(0, np.int8)
    ls4 = L[q].astype(np.int64)
    pure4 = (arrays.pure[q] != 0) & (ls4 > 0)
    nzr = nz.reshape(8, nq)
    nonzero = np.zeros(nq, np.int8)
    for k in range(4):
        nonzero |= nzr[k] | np.where(ls4[:, k] > 0, nzr[4 + k], 0).astype(np.int8)
    sph = {}
    # quartets of one class (same l and pure flag on each centre) share block
    # shapes: each class is combined in one batch, quartet index leading
    key = (ls4 * 2 + pure4) @ np.array([512, 64, 8, 1])
    for kv in np.unique(key):
        idx = np.nonzero(key == kv)[0]
        m = len(idx)
        ls = [int(x) for x in ls4[idx[0]]]
        pure = [bool(x) for x in pure4[idx[0]]]
        nc = [(l + 1) * (l + 2) // 2 for l in ls]
        osz = int(size[idx[0]])
        dst = out_off[idx][:, None] + np.arange(osz)

        def gather(j, dims):
            n = int(np.prod(dims))
            return vals[offs[j][:, None] + np.arange(n)].reshape([m] + dims)

        for k in range(4):
            dims_p = list(nc)
            dims_p[k] = (ls[k] + 2) * (ls[k] + 3) // 2
            ip = gather(k * nq + idx, dims_p)
            im = None
            if ls[k] > 0:
                dims_m = list(nc)
                dims_m[k] = ls[k] * (ls[k] + 1) // 2
                im = gather((4 + k) * nq + idx, dims_m)
            comps = _cart(ls[k])
            for d in range(3):
                up = [_cart_index(*[c[t] + (t == d) for t in range(3)]) for c in comps]
                val = np.take(ip, up, axis=k + 1)
                if im is not None:
                    fac = np.array([c[d] for c in comps], float)
                    dn = [_cart_index(*[c[t] - (t == d) for t in range(3)]) if c[d] > 0 else 0 for c in comps]
                    shape = [1] * 5
                    shape[k + 1] = len(comps)
                    val = val - fac.reshape(shape) * np.take(im, dn, axis=k + 1)
                for kk in range(4):  # solid harmonics on the pure centres
                    if pure[kk]:
                        if ls[kk] not in sph:
                            sph[ls[kk]] = _sph_matrix(ls[kk])
                        val = np.moveaxis(np.tensordot(sph[ls[kk]], val, axes=([1], [kk + 1])), 0, kk + 1)
                out[3 * k + d, dst] = val.reshape(m, osz)
    return out[:, :total], out_off, nonzero


def fock_deriv1(ctx, arrays: BasisArrays, shell2atom, natoms, D, precision=np.finfo(float).eps,
                screening=SCREEN_ORIGINAL):
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
    two-electron part of the energy gradient, d tr(D G(D)) / dx at fixed D."""
    ns = arrays.nshell
    sizes = arrays.sizes().astype(np.int64)
    bf = np.concatenate([[0], np.cumsum(sizes)])
    nbf = int(bf[-1])
    D = np.asarray(D, float)
    quartets = [(s1, s2, s3, s4) for s1 in range(ns) for s2 in range(s1 + 1) for s3 in range(s1 + 1)
                for s4 in range((s2 if s1 == s3 else s3) + 1)]
    q = np.array(quartets, np.int32).reshape(-1, 4)
    vals, off, nz = eri_deriv1(ctx, arrays, q, precision, screening)
    atom = np.asarray(shell2atom, np.int64)
    G = np.zeros((3 * natoms, nbf, nbf))
    for i, (s1, s2, s3, s4) in enumerate(quartets):
        if not nz[i]:
            continue
        n = [int(sizes[s]) for s in (s1, s2, s3, s4)]
        r = [slice(bf[s], bf[s] + sizes[s]) for s in (s1, s2, s3, s4)]
        deg = (1.0 if s1 == s2 else 2.0) * (1.0 if s3 == s4 else 2.0) * \
            (1.0 if (s1 == s3 and s2 == s4) else 2.0)
        blk = vals[:, off[i]:off[i] + int(np.prod(n))].reshape([12] + n)
        coord = np.array([3 * atom[(s1, s2, s3, s4)[c // 3]] + c % 3 for c in range(12)])
        uc, inv = np.unique(coord, return_inverse=True)
        V = np.zeros([len(uc)] + n)
        np.add.at(V, inv, blk)
        V *= deg
        G[uc, r[0], r[1]] += np.einsum("xabcd,cd->xab", V, D[r[2], r[3]])
        G[uc, r[2], r[3]] += np.einsum("xabcd,ab->xcd", V, D[r[0], r[1]])
        G[uc, r[0], r[2]] -= 0.25 * np.einsum("xabcd,bd->xac", V, D[r[1], r[3]])
        G[uc, r[1], r[3]] -= 0.25 * np.einsum("xabcd,ac->xbd", V, D[r[0], r[2]])
        G[uc, r[0], r[3]] -= 0.25 * np.einsum("xabcd,bc->xad", V, D[r[1], r[2]])
        G[uc, r[1], r[2]] -= 0.25 * np.einsum("xabcd,ad->xbc", V, D[r[0], r[3]])
    return 0.5 * (G + G.transpose(0, 2, 1))
