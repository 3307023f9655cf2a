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
    nonzero = np.zeros(nq, np.int8)
    sph = {}
    for i in range(nq):
        ls = [int(L[s]) for s in q[i]]
        nc = [(l + 1) * (l + 2) // 2 for l in ls]
        blocks = []
        for k in range(4):
            dims_p = list(nc)
            dims_p[k] = (ls[k] + 2) * (ls[k] + 3) // 2
            j = k * nq + i
            ip = vals[offs[j]:offs[j] + int(np.prod(dims_p))].reshape(dims_p)
            if ls[k] > 0:
                dims_m = list(nc)
                dims_m[k] = ls[k] * (ls[k] + 1) // 2
                j2 = (4 + k) * nq + i
                im = vals[offs[j2]:offs[j2] + int(np.prod(dims_m))].reshape(dims_m)
            else:
                im = None
            nonzero[i] |= nz[k * nq + i] | (nz[(4 + k) * nq + i] if ls[k] > 0 else 0)
            comps = _cart(ls[k])
            for d in range(3):
                up = [_cart_index(*[c[t] + (t == d) for t in range(3)]) for c in comps]
                val = np.take(ip, up, axis=k)
                if im is not None:
                    fac = np.array([c[d] for c in comps], float)
                    dn = [_cart_index(*[c[t] - (t == d) for t in range(3)]) if c[d] > 0 else 0 for c in comps]
                    shape = [1, 1, 1, 1]
                    shape[k] = len(comps)
                    val = val - fac.reshape(shape) * np.take(im, dn, axis=k)
                blocks.append((3 * k + d, val))
        for dd, val in blocks:
            for k in range(4):  # solid harmonics on the pure centres
                s = q[i, k]
                if arrays.pure[s] and ls[k] > 0:
                    if ls[k] not in sph:
                        sph[ls[k]] = _sph_matrix(ls[k])
                    val = np.moveaxis(np.tensordot(sph[ls[k]], val, axes=([1], [k])), 0, k)
            out[dd, out_off[i]:out_off[i] + size[i]] = val.reshape(-1)
    return out[:, :total], out_off, nonzero
