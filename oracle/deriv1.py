"""First geometric derivatives of the Coulomb ERIs, restated on the CPU --
test infrastructure: the checker of the device derivative plan
(l2g_eri_deriv1_plan_*), never called by the product.

libint's Engine(Operator::coulomb, max_nprim, max_l, deriv_order=1) returns,
per shell quartet, 12 shell sets d/dA_x, d/dA_y, d/dA_z, d/dB_x, ..., d/dD_z
for the centres of the shells in caller order (CartesianDerivIterator<4>,
include/libint2/util/intpart_iter.h:89-103, 161-182), each row-major in
caller order like the integrals themselves.  Differentiating a Cartesian
Gaussian with respect to its centre,

    d/dA_i  x^a_i e^{-alpha r^2} = 2 alpha x^(a_i+1) e^{...} - a_i x^(a_i-1) e^{...},

turns each derivative shell set into ERIs of the shell raised by one (the
contraction coefficients times 2 alpha) and lowered by one.  Here those ERIs
come from the oracle's restatement of Engine::compute2; the combination is
done per Cartesian component and pure centres are transformed to solid
harmonics afterwards (the transform commutes with the combination).
Pinned by Richardson-extrapolated central differences of the oracle's
integrals in the shell centres and by translational invariance
(tests/test_deriv.py), and end to end by the reference's 2-body force golden
(tests/test_hf_augccpvdz.py).
"""
from __future__ import annotations

import numpy as np

from . import ORIGINAL, OBasis, eri_quartets, solidharm_coeff


def _cart(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def _cart_index(x, y, z):
    l = x + y + z
    return (l - x) * (l - x + 1) // 2 + (l - x - y)


def _sph_matrix(l):
    """Rows m = -l..l (Standard ordering), columns Cartesian."""
    return np.array([[solidharm_coeff(l, m, *c) for c in _cart(l)] for m in range(-l, l + 1)])


def augmented(b):
    """Cartesian copies of the shells, then the raised (l+1, coefficients
    2 alpha c) and the lowered (l-1, same coefficients; an unused s copy for
    l = 0) shells."""
    b = OBasis.of(b)
    n = len(b.l)
    off = np.concatenate([[0], np.cumsum(b.nprim)])
    ls, pure, nprim, al, co, org = [], [], [], [], [], []
    for kind in range(3):
        for s in range(n):
            a = b.alpha[off[s]:off[s + 1]]
            c = b.coeff[off[s]:off[s + 1]]
            l = int(b.l[s])
            if kind == 1:
                l, c = l + 1, 2.0 * a * c
            elif kind == 2:
                l = max(l - 1, 0)
            ls.append(l)
            pure.append(0)
            nprim.append(len(a))
            al.append(a)
            co.append(c)
            org.append(b.origin[3 * s:3 * s + 3])
    return OBasis(np.array(ls, np.int32), np.array(pure, np.int32), np.array(nprim, np.int32),
                  np.concatenate(al), np.concatenate(co), np.concatenate(org))


def eri_deriv1(basis, quartets, precision=np.finfo(float).eps, screening=ORIGINAL):
    """First derivatives of (s0 s1|s2 s3) for each quartet (caller order).

    Returns (values, offsets, nonzero): values[d] holds the 12 derivative
    shell sets d = 3 * centre + axis back to back, quartet q's set at
    values[d][offsets[q]: offsets[q] + size_q]; nonzero[q] is 0 when every
    primitive of the raised and lowered quartets was screened out."""
    b = OBasis.of(basis)
    q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
    n = len(b.l)
    aug = augmented(b)
    big = [q + np.eye(4, dtype=np.int32)[k] * n for k in range(4)] + \
          [q + np.eye(4, dtype=np.int32)[k] * 2 * n for k in range(4)]
    allq = np.concatenate(big) if len(q) else np.zeros((0, 4), np.int32)
    vals, offs, nz = eri_quartets(aug, allq, precision=precision, screening=screening)
    nq = len(q)
    L = b.l.astype(np.int64)
    sizes = np.where(b.pure != 0, 2 * L + 1, (L + 1) * (L + 2) // 2)
    size = np.ones(nq, np.int64)
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
        for k in range(4):
            dims_p = list(nc)
            dims_p[k] = (ls[k] + 2) * (ls[k] + 3) // 2
            j = k * nq + i
            ip = vals[offs[j]:offs[j] + int(np.prod(dims_p))].reshape(dims_p)
            im = None
            if ls[k] > 0:
                dims_m = list(nc)
                dims_m[k] = ls[k] * (ls[k] + 1) // 2
                j2 = (4 + k) * nq + i
                im = vals[offs[j2]:offs[j2] + int(np.prod(dims_m))].reshape(dims_m)
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
                for kk in range(4):  # solid harmonics on the pure centres
                    if b.pure[q[i, kk]] and ls[kk] > 0:
                        if ls[kk] not in sph:
                            sph[ls[kk]] = _sph_matrix(ls[kk])
                        val = np.moveaxis(np.tensordot(sph[ls[kk]], val, axes=([1], [kk])), 0, kk)
                out[3 * k + d, out_off[i]:out_off[i] + size[i]] = val.reshape(-1)
    return out[:, :total], out_off, nonzero
