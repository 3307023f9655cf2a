"""One-body integrals over contracted Cartesian shells: overlap, kinetic
energy and nuclear attraction -- test infrastructure for the end-to-end
Hartree-Fock check (the reference computes them with Engine(Operator::overlap
/ kinetic / nuclear), include/libint2/engine.impl.h:181-561).

Restated with the McMurchie-Davidson Hermite expansion (Helgaker, Jorgensen,
Olsen, "Molecular Electronic-Structure Theory", sections 9.5, 9.7, 9.9):
E^{ij}_t coefficients of the Gaussian product, overlap and kinetic energy
from the 1-D overlaps, nuclear attraction from the Hermite Coulomb integrals
R_{tuv} with the Boys function F_n(T) = gamma(n+1/2, T) / (2 T^(n+1/2)).
Shell coefficients are used as given (libint's Shell already holds the
primitive normalisation of the x^l component); the functions of a shell are
in libint's Cartesian order (x descending, then y descending).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import gammainc, gamma as Gamma


def cart(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def boys(nmax, T):
    out = np.empty(nmax + 1)
    for n in range(nmax + 1):
        if T < 1e-13:
            out[n] = 1.0 / (2 * n + 1) - T / (2 * n + 3)
        else:
            a = n + 0.5
            out[n] = 0.5 * T ** (-a) * Gamma(a) * gammainc(a, T)
    return out


def hermite_E(i, j, a, b, XAB):
    """E[i][j][t] for exponents a, b and XAB = A_x - B_x."""
    p = a + b
    mu = a * b / p
    XPA = -b / p * XAB
    XPB = a / p * XAB
    E = np.zeros((i + 1, j + 1, i + j + 2))
    E[0, 0, 0] = math.exp(-mu * XAB * XAB)
    for ii in range(i + 1):
        for jj in range(j + 1):
            if ii == 0 and jj == 0:
                continue
            for t in range(ii + jj + 1):
                if ii > 0:
                    v = XPA * E[ii - 1, jj, t]
                    if t > 0:
                        v += E[ii - 1, jj, t - 1] / (2 * p)
                    v += (t + 1) * E[ii - 1, jj, t + 1]
                else:
                    v = XPB * E[ii, jj - 1, t]
                    if t > 0:
                        v += E[ii, jj - 1, t - 1] / (2 * p)
                    v += (t + 1) * E[ii, jj - 1, t + 1]
                E[ii, jj, t] = v
    return E


def hermite_R(tmax, p, PC):
    """R^0_{tuv} for t+u+v <= tmax."""
    T = p * float(np.dot(PC, PC))
    F = boys(tmax, T)
    S = tmax + 1
    R = np.zeros((tmax + 1, S, S, S))
    for n in range(tmax + 1):
        R[n, 0, 0, 0] = (-2 * p) ** n * F[n]
    for n in range(tmax - 1, -1, -1):
        top = tmax - n
        for t in range(top + 1):
            for u in range(top + 1 - t):
                for v in range(top + 1 - t - u):
                    if t + u + v == 0:
                        continue
                    if t > 0:
                        val = PC[0] * R[n + 1, t - 1, u, v] + ((t - 1) * R[n + 1, t - 2, u, v] if t > 1 else 0)
                    elif u > 0:
                        val = PC[1] * R[n + 1, t, u - 1, v] + ((u - 1) * R[n + 1, t, u - 2, v] if u > 1 else 0)
                    else:
                        val = PC[2] * R[n + 1, t, u, v - 1] + ((v - 1) * R[n + 1, t, u, v - 2] if v > 1 else 0)
                    R[n, t, u, v] = val
    return R[0]


def _prim_block(la, lb, a, b, A, B, charges):
    """Overlap, kinetic and nuclear-attraction blocks of two primitives."""
    p = a + b
    P = (a * A + b * B) / p
    Es = [hermite_E(la + 2, lb + 2, a, b, A[k] - B[k]) for k in range(3)]
    ca, cb = cart(la), cart(lb)
    S = np.zeros((len(ca), len(cb)))
    Tk = np.zeros_like(S)
    V = np.zeros_like(S)
    s1 = math.sqrt(math.pi / p)

    def ov(k, i, j):
        return Es[k][i, j, 0] * s1 if i >= 0 and j >= 0 else 0.0

    Rs = [(Z, hermite_R(la + lb, p, P - C)) for Z, C in charges]
    for ia, ea in enumerate(ca):
        for ib, eb in enumerate(cb):
            sx = [ov(k, ea[k], eb[k]) for k in range(3)]
            S[ia, ib] = sx[0] * sx[1] * sx[2]
            # -1/2 d^2/dx^2 on the ket, per axis
            t = 0.0
            for k in range(3):
                j = eb[k]
                tk = b * (2 * j + 1) * ov(k, ea[k], j) - 2 * b * b * ov(k, ea[k], j + 2)
                if j >= 2:
                    tk -= 0.5 * j * (j - 1) * ov(k, ea[k], j - 2)
                t += tk * sx[(k + 1) % 3] * sx[(k + 2) % 3]
            Tk[ia, ib] = t
            v = 0.0
            for Z, R in Rs:
                acc = 0.0
                for tt in range(ea[0] + eb[0] + 1):
                    for uu in range(ea[1] + eb[1] + 1):
                        for vv in range(ea[2] + eb[2] + 1):
                            acc += Es[0][ea[0], eb[0], tt] * Es[1][ea[1], eb[1], uu] * \
                                Es[2][ea[2], eb[2], vv] * R[tt, uu, vv]
                v -= Z * acc
            V[ia, ib] = v * 2 * math.pi / p
    return S, Tk, V


def one_body(shells, charges):
    """shells: list of (l, alphas, coeffs, origin) Cartesian; charges: list
    of (Z, position).  Returns S, T, V over all functions."""
    sizes = [(l + 1) * (l + 2) // 2 for l, _, _, _ in shells]
    off = np.concatenate([[0], np.cumsum(sizes)])
    n = int(off[-1])
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    for i, (la, aa, ca, A) in enumerate(shells):
        for j, (lb, ab, cb, B) in enumerate(shells):
            bs, bt, bv = 0, 0, 0
            for a, c1 in zip(aa, ca):
                for b, c2 in zip(ab, cb):
                    s, t, v = _prim_block(la, lb, a, b, np.asarray(A, float), np.asarray(B, float), charges)
                    bs = bs + c1 * c2 * s
                    bt = bt + c1 * c2 * t
                    bv = bv + c1 * c2 * v
            S[off[i]:off[i + 1], off[j]:off[j + 1]] = bs
            T[off[i]:off[i + 1], off[j]:off[j + 1]] = bt
            V[off[i]:off[i + 1], off[j]:off[j + 1]] = bv
    return S, T, V
