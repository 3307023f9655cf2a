"""Programs for the generic (table-driven) ERI kernel, libint_amd/csrc/l2g_generic.hip.

The register-resident kernels of gen_eri.py stop at (dd|dd): from f shells on,
a class's contracted [e0|f0] targets and recursion graph no longer fit a lane
(for (ff|ff): 5476 targets, 16-17k recursion values per primitive quartet).
The generic kernel evaluates one contracted quartet per 256-thread block with
everything in LDS, and reads the work of each class from data written here:

* OS VRR, one record per [e0|f0] node carrying the node's whole m-vector
  (m = 0 .. mlen-1).  Nodes are grouped into levels by |e| + |f|; level k only
  reads levels k-1 and k-2, so three LDS regions, used round-robin, hold the
  live levels.  The record of node (e, f) lowered along axis i:

    ket step (f != 0):  (e, f)^m = QC_i (e, f-1_i)^m + WQ_i (e, f-1_i)^{m+1}
        + (f_i - 1)/(2 eta) [(e, f-2_i)^m - rho/eta (e, f-2_i)^{m+1}]
        + e_i/(2 (zeta + eta)) (e-1_i, f-1_i)^{m+1}
    bra step (f == 0):  (e, 0)^m = PA_i (e-1_i, 0)^m + WP_i (e-1_i, 0)^{m+1}
        + (e_i - 1)/(2 zeta) [(e-2_i, 0)^m - rho/zeta (e-2_i, 0)^{m+1}]

  (Obara and Saika, J. Chem. Phys. 84, 3963 (1986); Head-Gordon and Pople,
  J. Chem. Phys. 89, 5777 (1988)), with the same axis choice as gen_eri.py.
  Nodes that are HGP targets ([e0|f0]^0, |e| in [la, la+lb], |f| in
  [lc, lc+ld]) add their m = 0 value to an LDS accumulator.

* HGP HRR programs for the ket (CD) and the bra (AB): (dst, src0, src1, axis)
  with dst = src0 + X_axis src1 over a per-row workspace whose first entries
  are the contracted [e0|f0] inputs, and the workspace index of every
  Cartesian output.

Records (int4): VRR node = {out | mlen << 15 | (kind * 3 + axis) << 19 |
n_a << 22 | n_c << 25 | kind << 28, X | Y << 16, Z | (target + 1) << 16, 0}
(offsets below 2^15 doubles, target 0 = none); HRR op = {dst, src0, src1,
axis}.  Offsets are in doubles of the block's LDS.
"""
try:  # as libint_amd.codegen.gen_generic
    from .gen_eri import cart, ncart, build_dir, sub1, hrr_program
except ImportError:  # as a script module next to gen_eri.py
    from gen_eri import cart, ncart, build_dir, sub1, hrr_program

ZERO_LEN = 16   # zero m-vector for absent terms
NCOEF = 20      # PA xyz, QC xyz, WP xyz, WQ xyz, oo2z, oo2e, roz, roe, oo2ze (padded)
LDS_DOUBLES = 160 * 1024 // 8 - 64  # 160 KB per CU on gfx950, minus the reduction scratch


def vrr_nodes(la, lb, lc, ld):
    """(e, f) -> m-vector length needed by the HGP targets."""
    need = {}

    def req(e, f, mlen):
        if need.get((e, f), 0) >= mlen:
            return
        need[(e, f)] = mlen
        if sum(f) > 0:
            i = build_dir(f)
            fm = sub1(f, i)
            req(e, fm, mlen + 1)
            if fm[i] > 0:
                req(e, sub1(fm, i), mlen + 1)
            if e[i] > 0:
                req(sub1(e, i), fm, mlen + 1)
        elif sum(e) > 0:
            i = build_dir(e)
            em = sub1(e, i)
            req(em, f, mlen + 1)
            if em[i] > 0:
                req(sub1(em, i), f, mlen + 1)

    ecomps = [c for l in range(la, la + lb + 1) for c in cart(l)]
    fcomps = [c for l in range(lc, lc + ld + 1) for c in cart(l)]
    for e in ecomps:
        for f in fcomps:
            req(e, f, 1)
    return need, ecomps, fcomps


def class_program(la, lb, lc, ld):
    L = la + lb + lc + ld
    need, ecomps, fcomps = vrr_nodes(la, lb, lc, ld)
    ne, nf = len(ecomps), len(fcomps)
    z = (0, 0, 0)
    levels = {}
    for (e, f) in sorted(need, key=lambda k: (sum(k[0]) + sum(k[1]), k[0][::-1], k[1][::-1])):
        levels.setdefault(sum(e) + sum(f), []).append((e, f))
    # LDS: [acc: ne nf contracted targets][zero][coefficients][3 level
    # regions]; after the primitive loop the regions are reused
    # strides are odd numbers of doubles: lanes of a wave walk neighbouring
    # m-vectors / rows in lockstep, and an even stride folds them onto few
    # of the LDS banks
    def odd(n):
        return n | 1

    accs = odd(nf)
    region = max(sum(odd(need[n]) for n in levels[k]) for k in levels)
    acc = 0
    zero = ne * accs
    coef = zero + ZERO_LEN
    reg0 = coef + NCOEF
    off = {}
    for k, nodes in levels.items():
        o = reg0 + (k % 3) * region
        for n in nodes:
            off[n] = o
            o += odd(need[n])
    eidx = {e: i for i, e in enumerate(ecomps)}
    fidx = {f: j for j, f in enumerate(fcomps)}

    def target(e, f):
        return eidx[e] * accs + fidx[f] if (e in eidx and f in fidx) else -1

    recs = []
    lev_first = []
    for k in range(1, L + 1):
        lev_first.append(len(recs))
        for (e, f) in levels.get(k, []):
            if sum(f) > 0:
                kind, i = 1, build_dir(f)
                x = (e, sub1(f, i))
                y = (e, sub1(sub1(f, i), i)) if f[i] >= 2 else None
                zz = (sub1(e, i), sub1(f, i)) if e[i] > 0 else None
                na, nc = f[i] - 1, e[i]
            else:
                kind, i = 0, build_dir(e)
                x = (sub1(e, i), f)
                y = (sub1(sub1(e, i), i), f) if e[i] >= 2 else None
                zz = None
                na, nc = e[i] - 1, 0
            o, ox = off[(e, f)], off[x]
            oy, oz = off[y] if y else zero, off[zz] if zz else zero
            tg = target(e, f) + 1
            assert max(o, ox, oy, oz) < (1 << 15) and tg < (1 << 16) and need[(e, f)] < 16
            recs.append((o | (need[(e, f)] << 15) | ((kind * 3 + i) << 19) | (na << 22) | (nc << 25) | (kind << 28),
                         ox | (oy << 16), oz | (tg << 16), 0))
    lev_first.append(len(recs))
    base = off[(z, z)]
    m0len = need[(z, z)]
    targ0 = target(z, z)

    def hrr(l1, l2, comps):
        ops, outs = hrr_program(l1, l2)
        ws = {(c, z): j for j, c in enumerate(comps)}
        prog = []
        for n, (key, s0, s1, i) in enumerate(ops):
            ws[key] = len(comps) + n
            prog.append((ws[key], ws[s0], ws[s1], i))
        return prog, [ws[k] for k in outs], len(comps) + len(ops)

    hk, hk_out, hk_nws = hrr(lc, ld, fcomps)
    hb, hb_out, hb_nws = hrr(la, lb, ecomps)
    # after the primitive loop, ket phase: acc is read, Y = ket-transformed
    # rows [ne][ncart(lc) ncart(ld)], then chk row workspaces of hk_nws + ncc
    # doubles.  Bra phase: Y is read; the digest accumulators (J12, K13, K14,
    # K24, K23 of the Fock kernels) and chb column workspaces of hb_nws + nab
    # doubles follow Y.
    ncc = ncart(lc) * ncart(ld)
    nab = ncart(la) * ncart(lb)
    ys = odd(ncc)
    lds_vrr = reg0 + 3 * region
    y0 = acc + ne * accs
    w0 = y0 + ne * ys
    wrow = odd(hk_nws + ncc)
    # rows / columns per pass: as many as fit under the larger of the
    # recursion's own footprint and a quarter of the CU's LDS, so that small
    # classes keep several blocks per CU, unless that leaves fewer than 32
    # rows / columns per pass; then as many as the CU's LDS holds
    cap = max(lds_vrr, LDS_DOUBLES // 4)

    def per_pass(n, start, size):
        capped = (cap - start) // size
        return min(n, 256, capped if capped >= min(n, 32) else (LDS_DOUBLES - start) // size)

    chk = per_pass(ne, w0, wrow)
    # digest accumulators: a low and a high word each, for the fixed-point
    # (deterministic) accumulation; the FP64 accumulation uses the first half
    dg = 2 * (nab + (ncart(la) + ncart(lb)) * (ncart(lc) + ncart(ld)))
    wcol = odd(hb_nws + nab)
    wb0 = w0 + dg
    chb = per_pass(ncc, wb0, wcol)
    if chk < 1 or chb < 1:
        raise ValueError(f"class ({la}{lb}|{lc}{ld}) does not fit {LDS_DOUBLES} doubles of LDS")
    lds = max(lds_vrr, w0 + chk * wrow, wb0 + chb * wcol)
    if lds > LDS_DOUBLES:
        raise ValueError(f"class ({la}{lb}|{lc}{ld}) needs {lds} doubles of LDS")
    f0 = w0  # digest accumulators
    # two primitive quartets per pass of the recursion when a second copy of
    # the [zero][coefficients][level regions] block, pstride doubles further
    # on, fits the LDS the class already takes: half the barriers per
    # primitive quartet, and each thread's nodes of both in flight together
    pstride = reg0 + 3 * region - zero
    dual = 1 if zero + 2 * pstride <= lds else 0
    return dict(la=la, lb=lb, lc=lc, ld=ld, L=L, ne=ne, nf=nf, m0len=m0len, targ0=targ0, dual=dual, pstride=pstride,
                nvals=sum(need.values()),
                base=base, zero=zero, coef=coef, acc=acc, y0=y0, w0=w0, wrow=wrow, chk=chk,
                f0=f0, wb0=wb0, wcol=wcol, chb=chb, lds=lds, recs=recs, lev_first=lev_first,
                hk=hk, hk_out=hk_out, hk_nws=hk_nws, hb=hb, hb_out=hb_out, hb_nws=hb_nws,
                accs=accs, ys=ys,
                diag=[i * accs + i for i in range(ne)] if (la, lb) == (lc, ld) else [])


FIELDS = ["la", "lb", "lc", "ld", "L", "ne", "nf", "m0len", "targ0", "base", "zero", "coef", "acc",
          "y0", "w0", "wrow", "chk", "f0", "wb0", "wcol", "chb", "lds", "vrr_first", "lev_first",
          "hk_first", "hk_nops", "hk_out", "hk_nws", "hb_first", "hb_nops", "hb_out", "hb_nws",
          "diag_first", "ndiag", "nvals", "accs", "ys", "dual", "pstride"]


def emit_tables(classes, lmax):
    """C++ source with the programs of `classes` (list of (la,lb,lc,ld)),
    the canonical classes up to lmax."""
    prog = []   # int4 records
    idx = []    # ints: level starts, HRR outputs, diagonal targets
    rows = []
    for c in classes:
        p = class_program(*c)
        d = {k: p[k] for k in FIELDS if k in p}
        d["vrr_first"] = len(prog) // 1
        prog.extend(p["recs"])
        d["lev_first"] = len(idx)
        idx.extend(p["lev_first"])
        d["hk_first"] = len(prog)
        d["hk_nops"] = len(p["hk"])
        prog.extend(p["hk"])
        d["hk_out"] = len(idx)
        idx.extend(p["hk_out"])
        d["hb_first"] = len(prog)
        d["hb_nops"] = len(p["hb"])
        prog.extend(p["hb"])
        d["hb_out"] = len(idx)
        idx.extend(p["hb_out"])
        d["diag_first"] = len(idx)
        d["ndiag"] = len(p["diag"])
        idx.extend(p["diag"])
        rows.append(d)
    out = ["// generated by libint_amd/codegen/gen_generic.py — do not edit",
           '#include "eri_device.h"', "",
           f"extern const int l2g_gen_generic_lmax = {lmax};",
           f"extern const int l2g_gen_ngeneric = {len(rows)};",
           "extern const L2gGenericClass l2g_gen_generic[] = {"]
    for d in rows:
        out.append("  {" + ", ".join(str(d[k]) for k in FIELDS) + "},")
    if not rows:
        out.append("  {" + ", ".join("0" for _ in FIELDS) + "},")
    out.append("};")
    prog.append((0, 0, 0, 0))  # padding: a clamped prefetch never reads past the array
    out.append(f"extern const int l2g_gen_generic_nprog = {len(prog)};")
    out.append("extern const int l2g_gen_generic_prog[] = {")
    for i in range(0, len(prog), 4):
        out.append("  " + ", ".join(f"{a},{b},{c_},{e}" for (a, b, c_, e) in prog[i:i + 4]) + ",")
    if not prog:
        out.append("  0, 0, 0, 0,")
    out.append("};")
    out.append(f"extern const int l2g_gen_generic_nidx = {len(idx)};")
    out.append("extern const int l2g_gen_generic_idx[] = {")
    for i in range(0, len(idx), 16):
        out.append("  " + ", ".join(str(v) for v in idx[i:i + 16]) + ",")
    if not idx:
        out.append("  0,")
    out.append("};")
    return "\n".join(out) + "\n"
