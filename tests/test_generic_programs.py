"""CPU check of the generic kernel's programs (codegen/gen_generic.py): a
numpy walk over the same records (VRR m-vector nodes in three rotating LDS
regions, target accumulation, ket and bra HRR programs) must reproduce the
oracle's Cartesian integrals for classes up to f.  This pins the tables the
GPU kernel reads; the kernel's own arithmetic is checked on the GPU."""
import numpy as np
import pytest

import libint_amd as la
import oracle
from libint_amd.codegen import gen_eri, gen_generic


def cart_shell(l, prims, origin):
    return la.Shell(l, prims, origin, pure=False)


def emulate(prog, shells, dual=False):
    """Contracted Cartesian (ab|cd) of the canonical class `prog` for four
    renormalised Cartesian shells, walking the program records.  dual: the
    kernel's two-primitive pass (prog["dual"] classes): primitive quartets
    in pairs, the second one's zero vector, coefficients and level regions
    prog["pstride"] doubles after the first's, both added per node in
    primitive order."""
    A, B, C, D = (np.array(s.O) for s in shells)
    la_, lb_, lc_, ld_ = prog["la"], prog["lb"], prog["lc"], prog["ld"]
    ne, nf, L = prog["ne"], prog["nf"], prog["L"]
    lds = np.zeros(prog["lds"] + 64)
    accs = prog["accs"]
    acc = np.zeros(ne * accs)
    rec = prog["recs"]
    lev = prog["lev_first"]

    def pairs(s1, s2, P1, P2):
        out = []
        AB = P1 - P2
        for a1, c1 in zip(s1.alpha, s1.coeff):
            for a2, c2 in zip(s2.alpha, s2.coeff):
                g = a1 + a2
                P = (a1 * P1 + a2 * P2) / g
                K = np.sqrt(2) * np.pi ** 1.25 * np.exp(-a1 * a2 / g * AB @ AB) / g * c1 * c2
                out.append((g, P, K))
        return out

    prims = []
    for (gp, P, Kb) in pairs(shells[0], shells[1], A, B):
        for (gq, Q, Kk) in pairs(shells[2], shells[3], C, D):
            rho = gp * gq / (gp + gq)
            PQ = P - Q
            T = rho * PQ @ PQ
            pf = Kb * Kk / np.sqrt(gp + gq)
            coef = np.zeros(20)
            coef[0:3], coef[3:6] = P - A, Q - C
            coef[6:9], coef[9:12] = -gq / (gp + gq) * PQ, gp / (gp + gq) * PQ
            coef[12], coef[13] = 0.5 / gp, 0.5 / gq
            coef[14], coef[15], coef[16] = gq / (gp + gq), gp / (gp + gq), 0.5 / (gp + gq)
            F = oracle.boys_reference(T, L)[0]
            prims.append((coef, pf * F[:prog["m0len"]]))
    S = prog["pstride"] if dual else 0
    step = 2 if dual else 1
    for i0 in range(0, len(prims), step):
        group = prims[i0:i0 + step]
        for s, (_, base) in enumerate(group):
            lds[prog["zero"] + s * S:prog["zero"] + s * S + 16] = 0
            lds[prog["base"] + s * S:prog["base"] + s * S + prog["m0len"]] = base
        for k in range(1, L + 1):
            for nd in range(lev[k - 1], lev[k]):
                x, y, z, _ = rec[nd]
                out, mlen, slot = x & 0x7FFF, (x >> 15) & 15, (x >> 19) & 7
                na, nc, kind = (x >> 22) & 7, (x >> 25) & 7, (x >> 28) & 1
                ox, oy, oz, tg = y & 0xFFFF, y >> 16, z & 0xFFFF, (z >> 16) - 1
                for s, (coef, _) in enumerate(group):
                    o = s * S
                    cA, cW = coef[slot], coef[6 + slot]
                    g, r, h = na * coef[12 + kind], coef[14 + kind], nc * coef[16]
                    for m in range(mlen):
                        lds[o + out + m] = (cA * lds[o + ox + m] + cW * lds[o + ox + m + 1]
                                            + g * (lds[o + oy + m] - r * lds[o + oy + m + 1])
                                            + h * lds[o + oz + m + 1])
                if tg >= 0:
                    for s in range(len(group)):
                        acc[tg] += lds[s * S + out]
        if prog["targ0"] >= 0:
            for s in range(len(group)):
                acc[prog["targ0"]] += lds[prog["base"] + s * S]
    CD, AB = C - D, A - B
    ncc = gen_eri.ncart(lc_) * gen_eri.ncart(ld_)
    Y = np.zeros((ne, ncc))
    for ie in range(ne):
        ws = np.zeros(prog["hk_nws"])
        ws[:nf] = acc[ie * accs:ie * accs + nf]
        for (d, s0, s1, ax) in prog["hk"]:
            ws[d] = ws[s0] + CD[ax] * ws[s1]
        Y[ie] = ws[prog["hk_out"]]
    nab = gen_eri.ncart(la_) * gen_eri.ncart(lb_)
    out = np.zeros((nab, ncc))
    for col in range(ncc):
        ws = np.zeros(prog["hb_nws"])
        ws[:ne] = Y[:, col]
        for (d, s0, s1, ax) in prog["hb"]:
            ws[d] = ws[s0] + AB[ax] * ws[s1]
        out[:, col] = ws[prog["hb_out"]]
    return out.reshape(-1)


@pytest.mark.parametrize("cls", [(0, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1), (2, 0, 2, 1), (3, 0, 1, 0),
                                 (2, 1, 3, 0), (3, 1, 2, 2), (3, 2, 3, 1)])
def test_generic_program_matches_oracle(cls):
    la_, lb_, lc_, ld_ = cls
    prog = gen_generic.class_program(*cls)
    rng = np.random.default_rng(sum(cls) + 17)
    shells = []
    for l in cls:
        K = int(rng.integers(1, 3))
        shells.append(cart_shell(l, list(zip(rng.uniform(0.3, 2.0, K), rng.uniform(0.2, 1.0, K))),
                                 tuple(rng.uniform(-0.8, 0.8, 3))))
    v = emulate(prog, shells)
    arrays = la.BasisArrays.from_shells(shells)
    o, _, _ = oracle.eri_quartets(arrays, [[0, 1, 2, 3]], precision=0.0)
    assert v.shape == o.shape
    d = np.abs(v - o)
    assert (d <= 1e-12 + 1e-10 * np.abs(o)).all(), d.max()


@pytest.mark.parametrize("cls", [(2, 2, 3, 3), (3, 1, 3, 3), (3, 2, 3, 3)])
def test_generic_dual_pass(cls):
    """The kernel's two-primitive pass on classes whose LDS holds it: the
    second block stays inside the class's LDS and off the first block and
    the accumulators, and walking the records in pairs (with an odd count,
    so the last pass has one) gives the single-pass values bit for bit."""
    prog = gen_generic.class_program(*cls)
    assert prog["dual"] == 1
    zero, ps = prog["zero"], prog["pstride"]
    assert prog["acc"] + prog["ne"] * prog["accs"] <= zero and zero + 2 * ps <= prog["lds"]
    for (x, y, z, _) in prog["recs"]:
        mlen = (x >> 15) & 15
        for o, n in ((x & 0x7FFF, mlen), (y & 0xFFFF, mlen + 1), (y >> 16, mlen + 1), (z & 0xFFFF, mlen + 1)):
            assert zero <= o and o + n <= zero + ps
    rng = np.random.default_rng(sum(cls))
    ks = [1, 3, 1, 1]  # 3 primitive quartets
    shells = [cart_shell(l, list(zip(rng.uniform(0.3, 2.0, k), rng.uniform(0.2, 1.0, k))),
                         tuple(rng.uniform(-0.8, 0.8, 3))) for l, k in zip(cls, ks)]
    v1 = emulate(prog, shells)
    v2 = emulate(prog, shells, dual=True)
    assert np.array_equal(v1, v2)


def test_generic_programs_fit_lds():
    for c in gen_eri.canonical_classes(3):
        p = gen_generic.class_program(*c)
        assert p["lds"] <= gen_generic.LDS_DOUBLES
        assert p["chk"] >= 1 and p["chb"] >= 1


def test_generic_levels_write_apart_from_their_reads():
    """Within a level, no record's output range meets any record's operand
    ranges (X, Y, Z with one more m): the kernel's threads write a level
    while others read, and each thread carries X[m], Y[m] over its m-loop in
    registers, both sound only when a level's writes and reads are apart."""
    for c in gen_eri.canonical_classes(3):
        p = gen_generic.class_program(*c)
        recs, lev = p["recs"], p["lev_first"]
        for k in range(len(lev) - 1):
            w = np.zeros(p["lds"], bool)
            r = np.zeros(p["lds"], bool)
            for (x, y, z, _) in recs[lev[k]:lev[k + 1]]:
                mlen = (x >> 15) & 15
                w[x & 0x7FFF:(x & 0x7FFF) + mlen] = True
                for o in (y & 0xFFFF, y >> 16, z & 0xFFFF):
                    r[o:o + mlen + 1] = True
            assert not (w & r).any(), (c, k)
