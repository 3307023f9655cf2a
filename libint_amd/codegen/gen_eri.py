"""Generator of the gfx950 ERI kernels (one kernel per canonical class).

Each generated kernel evaluates contracted (ab|cd) Coulomb integrals for one
canonical angular-momentum class (la>=lb, lc>=ld, la+lb<=lc+ld, the class
predicate of libint's generator, src/bin/libint/build_libint.cc:79-83) with
the Head-Gordon-Pople scheme libint uses by default
(src/bin/libint/strategy.cc:67-80: VRR on a and c, HRR on b and d):

  1. Obara-Saika VRR builds [e0|f0]^(0), e in [la, la+lb], f in [lc, lc+ld]
     (src/bin/libint/vrr_11_twoprep_11.h:152-224), accumulated over the
     primitive quartets (src/bin/libint/dg.cc:1177-1189);
  2. HRR (ab| <- (a+1 b-1| + AB (a b-1|, and the same on the ket with CD
     (src/bin/libint/hrr.h:290-330);
  3. Cartesian -> real solid harmonics with the coefficients of
     include/libint2/solidharmonics.h:114-180;
  4. a strided write that undoes the canonical reorder
     (include/libint2/engine.impl.h:1987-2067).

Mapping to the hardware: one lane owns one shell quartet and keeps every
recurrence intermediate in VGPRs.  The contracted target set of a large class
does not fit one lane's registers ((dd|dd): 961 contracted values plus ~330
live intermediates), so the targets are partitioned by their bra Cartesian
component into groups whose recursion sub-graph fits a register budget.  Each
group runs the whole primitive loop on its own (recomputing the shared part of
the recursion), transforms its rows through the ket HRR + ket spherical
transform, and parks them in a lane-private scratch column in HBM; a final
pass reads the rows back column by column for the bra HRR, bra spherical
transform and the permuted store.  Classes whose targets fit in one group keep
everything in registers.

Run as a script: writes the kernels and the dispatch table into the output
directory given on the command line.
"""
from __future__ import annotations

import math
import os
import sys
from dataclasses import dataclass, field

# ---------------------------------------------------------------------------
# shell conventions
# ---------------------------------------------------------------------------


def cart(l):
    """Cartesian components of level l, libint standard order
    (include/libint2/cgshell_ordering.h:125-130)."""
    return [(i, j, l - i - j) for i in range(l, -1, -1) for j in range(l - i, -1, -1)]


def ncart(l):
    return (l + 1) * (l + 2) // 2


def cart_index(t):
    x, y, z = t
    l = x + y + z
    return ((l - x + 1) * (l - x)) // 2 + l - x - y


def _fac(n):
    return float(math.factorial(n)) if n >= 0 else 1.0


def _df_km1(k):
    r = 1.0
    i = k - 1
    while i > 1:
        r *= i
        i -= 2
    return r


def _binom(n, k):
    if k < 0 or k > n:
        return 0.0
    return float(math.comb(n, k))


def _parity(i):
    return -1 if i % 2 else 1


def sh_coeff(l, m, lx, ly, lz):
    """Coefficient of Cartesian (lx,ly,lz) in real solid harmonic (l,m),
    IJQC 54, 83 (1995) eq. 15 — the formula libint evaluates in
    include/libint2/solidharmonics.h:114-180."""
    abs_m = abs(m)
    if (lx + ly - abs_m) % 2:
        return 0.0
    j = (lx + ly - abs_m) // 2
    if j < 0:
        return 0.0
    comp = 1 if m >= 0 else -1
    i = abs_m - lx
    if comp != _parity(abs(i)):
        return 0.0
    pfac = math.sqrt(((_fac(2 * lx) * _fac(2 * ly) * _fac(2 * lz)) / _fac(2 * l)) *
                     (_fac(l - abs_m) / _fac(l)) * (1.0 / _fac(l + abs_m)) *
                     (1.0 / (_fac(lx) * _fac(ly) * _fac(lz))))
    pfac /= float(1 << l)
    # C integer division truncates toward zero
    pfac *= _parity(int((i - 1) / 2)) if m < 0 else _parity(int(i / 2))
    s = 0.0
    for ii in range(j, (l - abs_m) // 2 + 1):
        p1 = _binom(l, ii) * _binom(ii, j)
        p1 *= (_parity(ii) * _fac(2 * (l - ii))) / _fac(l - abs_m - 2 * ii)
        s1 = 0.0
        for k in range(max((lx - abs_m) // 2, 0), min(j, lx // 2) + 1):
            if lx - 2 * k <= abs_m:
                s1 += _binom(j, k) * _binom(abs_m, lx - 2 * k) * _parity(k)
        s += p1 * s1
    s *= math.sqrt(_df_km1(2 * l) / (_df_km1(2 * lx) * _df_km1(2 * ly) * _df_km1(2 * lz)))
    return pfac * s if m == 0 else math.sqrt(2.0) * pfac * s


def sh_rows(l):
    """Standard-order (m=-l..l) rows: list of [(cart index, coeff)]."""
    rows = []
    for m in range(-l, l + 1):
        r = []
        for ic, c in enumerate(cart(l)):
            v = sh_coeff(l, m, *c)
            if v != 0.0:
                r.append((ic, v))
        rows.append(r)
    return rows


def build_dir(t):
    """Axis used to lower a Cartesian triple: the one with the largest quantum
    number (ties -> x, y, z order)."""
    best = 0
    for i in (1, 2):
        if t[i] > t[best]:
            best = i
    return best


def sub1(t, i):
    u = list(t)
    u[i] -= 1
    return tuple(u)


def add1(t, i):
    u = list(t)
    u[i] += 1
    return tuple(u)


# ---------------------------------------------------------------------------
# Obara-Saika VRR graph
# ---------------------------------------------------------------------------


@dataclass
class VNode:
    key: tuple
    kind: str          # 'base' | 'bra' | 'ket'
    dir: int = -1
    deps: list = field(default_factory=list)   # [(coef_expr, node_key)]


class VrrGraph:
    def __init__(self):
        self.nodes: dict = {}

    def get(self, e, f, m):
        k = (e, f, m)
        if k in self.nodes:
            return k
        if sum(f) > 0:
            i = build_dir(f)
            f1 = sub1(f, i)
            n = VNode(k, 'ket', i)
            n.deps.append(('QC', self.get(e, f1, m)))
            n.deps.append(('WQ', self.get(e, f1, m + 1)))
            if f1[i] > 0:
                f2 = sub1(f1, i)
                n.deps.append((('f', f1[i]), self.get(e, f2, m)))
                n.deps.append((('fr', f1[i]), self.get(e, f2, m + 1)))
            if e[i] > 0:
                n.deps.append((('x', e[i]), self.get(sub1(e, i), f1, m + 1)))
        elif sum(e) > 0:
            i = build_dir(e)
            e1 = sub1(e, i)
            n = VNode(k, 'bra', i)
            n.deps.append(('PA', self.get(e1, f, m)))
            n.deps.append(('WP', self.get(e1, f, m + 1)))
            if e1[i] > 0:
                e2 = sub1(e1, i)
                n.deps.append((('e', e1[i]), self.get(e2, f, m)))
                n.deps.append((('er', e1[i]), self.get(e2, f, m + 1)))
        else:
            n = VNode(k, 'base')
        self.nodes[k] = n
        return k

    def order(self, targets):
        """DFS post-order over the sub-graph reaching `targets`; children are
        visited deepest-first, which keeps the live set small."""
        depth = {}

        def d(k):
            if k in depth:
                return depth[k]
            n = self.nodes[k]
            depth[k] = 1 + max((d(c) for _, c in n.deps), default=0)
            return depth[k]

        seen = set()
        out = []
        sys.setrecursionlimit(1000000)

        def visit(k):
            if k in seen:
                return
            n = self.nodes[k]
            for _, c in sorted(n.deps, key=lambda x: -d(x[1])):
                visit(c)
            seen.add(k)
            out.append(k)

        for t in targets:
            visit(t)
        return out


def peak_live(order, graph, targets):
    tset = set(targets)
    last = {}
    for idx, k in enumerate(order):
        for _, c in graph.nodes[k].deps:
            last[c] = idx
    cur = set()
    peak = 0
    for idx, k in enumerate(order):
        cur.add(k)
        for _, c in graph.nodes[k].deps:
            if last.get(c) == idx:
                cur.discard(c)
        if k not in last:
            cur.discard(k)
        peak = max(peak, len(cur))
    return peak


# ---------------------------------------------------------------------------
# class description
# ---------------------------------------------------------------------------


@dataclass
class ClassPlan:
    la: int
    lb: int
    lc: int
    ld: int
    ecomps: list
    fcomps: list
    groups: list            # list of (e indices, f indices) target blocks
    group_nodes: list       # node count per group
    group_peak: list

    def full_rows(self, g):
        return len(self.groups[g][1]) == len(self.fcomps)

    def row_last_group(self, ie):
        """index of the last group that contributes to bra row ie"""
        return max(g for g, (es, fs) in enumerate(self.groups) if ie in es)

    @property
    def split_rows(self):
        return any(not self.full_rows(g) for g in range(len(self.groups)))

    @property
    def L(self):
        return self.la + self.lb + self.lc + self.ld

    @property
    def name(self):
        return f"{self.la}{self.lb}{self.lc}{self.ld}"

    @property
    def ncol_cart(self):
        return ncart(self.lc) * ncart(self.ld)


def plan_class(la, lb, lc, ld, budget=96):
    """Partition the contracted targets [e0|f0] into blocks whose recursion
    sub-graph (live intermediates + accumulators) fits `budget` doubles:
    whole bra rows are merged while they fit; a bra row that alone exceeds
    the budget is cut into chunks of ket targets."""
    ecomps = [c for l in range(la, la + lb + 1) for c in cart(l)]
    fcomps = [c for l in range(lc, lc + ld + 1) for c in cart(l)]
    nf = len(fcomps)
    allf = list(range(nf))
    g = VrrGraph()
    order_e = sorted(range(len(ecomps)), key=lambda i: ecomps[i], reverse=True)
    order_f = sorted(range(nf), key=lambda j: fcomps[j], reverse=True)

    def cost(es, fs):
        tg = [g.get(ecomps[i], fcomps[j], 0) for i in es for j in fs]
        o = g.order(tg)
        return len(o), peak_live(o, g, tg) + len(tg)

    groups = []
    cur = []
    for i in order_e:
        if cost(cur + [i], allf)[1] <= budget:
            cur.append(i)
            continue
        if cur:
            groups.append((cur, allf))
            cur = []
        if cost([i], allf)[1] <= budget:
            cur = [i]
            continue
        # ket chunks keep the target order they were costed in: the
        # evaluation order, and so the live set, follows the target order
        fcur = []
        for j in order_f:
            if not fcur or cost([i], fcur + [j])[1] <= budget:
                fcur.append(j)
            else:
                groups.append(([i], fcur))
                fcur = [j]
        groups.append(([i], fcur))
    if cur:
        groups.append((cur, allf))
    gn, gp = [], []
    for es, fs in groups:
        n, pk = cost(es, fs)
        gn.append(n)
        gp.append(pk)
    return ClassPlan(la, lb, lc, ld, ecomps, fcomps, groups, gn, gp)


# ---------------------------------------------------------------------------
# code emission helpers
# ---------------------------------------------------------------------------

AX = 'xyz'


# arithmetic type of the recursion, HRR and transforms of the kernel being
# emitted ('double', or 'float' for the reduced-precision variants)
_R = 'double'
# attenuated (erf) variant: the core integrals s^(m+1/2) F_m(sT), s = w^2/(w^2+rho)
_ATT = False
# contraction on the matrix cores: each target's primitive sum accumulates
# through v_mfma_f64_4x4x4_4b_f64 with a one-hot B operand (see emit_class)
_MFMA = False
# Lane layout of v_mfma_f64_4x4x4_4b_f64 on gfx950 (4 blocks g of 4x4x4,
# probed on the device): A(g, i, k) sits in lane 4g + i + 16k, B(g, k, j) in
# lane 4g + j + 16k and D(g, i, j) in lane 16i + 4g + j.  With B(g, k, j) =
# [k == j] (MFMA_SELECTOR), D(g, i, j) += A(g, i, j): the value of lane
# 16j + 4g + i lands in the accumulator of lane 16i + 4g + j, a transpose of
# the lane bits [5:4] <-> [1:0].  Each product but one is an exact zero, so
# D = C + v is rounded once, as a plain add is; after the primitive loop one
# shuffle from lane MFMA_SOURCE returns every accumulator to its own lane (a
# wave sum needs none).
MFMA_SELECTOR = "((lane >> 4) == (lane & 3) ? 1.0 : 0.0)"
MFMA_SOURCE = "(((lane & 3) << 4) | (lane & 12) | (lane >> 4))"


def fmt(v):
    r = repr(float(v))
    if 'e' not in r and '.' not in r and 'inf' not in r and 'nan' not in r:
        r += '.0'
    return r + ('f' if _R == 'float' else '')


class Emitter:
    def __init__(self):
        self.lines = []
        self.ind = 0

    def __call__(self, s=''):
        self.lines.append('  ' * self.ind + s if s else '')

    def block(self, head):
        self(head + ' {')
        self.ind += 1

    def end(self, tail=''):
        self.ind -= 1
        self('}' + tail)

    def text(self):
        return '\n'.join(self.lines) + '\n'


def vrr_group_order(plan, graph, grp_idx):
    es, fs = plan.groups[grp_idx]
    targets = [graph.get(plan.ecomps[i], plan.fcomps[j], 0) for i in es for j in fs]
    return targets, graph.order(targets)


def vrr_coef_names(plan, graph, grp_idx):
    """Integer multiplicities used by the group's recursion, per kind."""
    targets, order = vrr_group_order(plan, graph, grp_idx)
    need = {'e': set(), 'f': set(), 'x': set()}
    for key in order:
        for c, _ in graph.nodes[key].deps:
            if isinstance(c, tuple) and c[0] in ('e', 'f', 'x'):
                need[c[0]].add(c[1])
    return need


def emit_vrr_coefs(E, need):
    """Per-primitive multiples k/(2 zeta), k rho/(2 zeta^2), ... so that every
    recursion term is a single fused multiply-add."""
    for k in sorted(need['e']):
        E(f"const {_R} ez{k} = {fmt(k)} * oo2z, ezr{k} = ez{k} * roz;")
    for k in sorted(need['f']):
        E(f"const {_R} fe{k} = {fmt(k)} * oo2e, fer{k} = fe{k} * roe;")
    for k in sorted(need['x']):
        E(f"const {_R} xz{k} = {fmt(k)} * oo2ze;")


def emit_vrr_group(E, plan, graph, grp_idx, acc_names, mid=None):
    """Emit the VRR sub-graph for group `grp_idx`; accumulate targets.
    `mid`, if given, is called halfway through the node list (the pipelined
    loop issues the next primitive's Boys-table load there)."""
    targets, order = vrr_group_order(plan, graph, grp_idx)
    name = {}
    tpos = {t: k for k, t in enumerate(targets)}
    half = len(order) // 2
    for k, key in enumerate(order):
        if mid is not None and k == half:
            mid()
        n = graph.nodes[key]
        v = f"v{k}"
        name[key] = v
        if n.kind == 'base':
            expr = f"Fv[{key[2]}]"
        elif n.kind == 'bra':
            i = AX[n.dir]
            expr = f"PA{i} * {name[n.deps[0][1]]} + WP{i} * {name[n.deps[1][1]]}"
            if len(n.deps) > 2:
                (ce, k2), (_, k3) = n.deps[2], n.deps[3]
                expr += f" + ez{ce[1]} * {name[k2]} - ezr{ce[1]} * {name[k3]}"
        else:
            i = AX[n.dir]
            expr = f"QC{i} * {name[n.deps[0][1]]} + WQ{i} * {name[n.deps[1][1]]}"
            rest = n.deps[2:]
            j = 0
            while j < len(rest):
                c, kk = rest[j]
                if c[0] == 'f':
                    expr += f" + fe{c[1]} * {name[kk]} - fer{c[1]} * {name[rest[j + 1][1]]}"
                    j += 2
                else:
                    expr += f" + xz{c[1]} * {name[kk]}"
                    j += 1
        E(f"const {_R} {v} = {expr};")
        if key in tpos:
            a = acc_names[tpos[key]]
            if _MFMA:
                E(f"{a} = __builtin_amdgcn_mfma_f64_4x4x4f64({v}, msel, {a}, 0, 0, 0);")
            else:
                E(f"{a} += {v};")
    if mid is not None and half >= len(order):
        mid()
    return len(order)


def hrr_program(l1, l2):
    """HRR from blocks e in [l1, l1+l2] to (a in l1, b in l2).
    Returns (ops, outputs): ops is a list of (dst_key, src0_key, src1_key, dir)
    in evaluation order; keys are (a, b) with b=(0,0,0) meaning an input."""
    nodes = {}
    order = []

    def get(a, b):
        k = (a, b)
        if k in nodes:
            return k
        if sum(b) == 0:
            nodes[k] = None
            return k
        i = build_dir(b)
        s0 = get(add1(a, i), sub1(b, i))
        s1 = get(a, sub1(b, i))
        nodes[k] = (s0, s1, i)
        order.append(k)
        return k

    outs = [get(a, b) for a in cart(l1) for b in cart(l2)]
    ops = [(k, nodes[k][0], nodes[k][1], nodes[k][2]) for k in order]
    return ops, outs


def emit_hrr(E, l1, l2, inp, prefix, dvec):
    """inp: dict e-triple -> C expression. Returns list of output expressions
    (a-major, b-minor, Cartesian order).  dvec: name of the distance vector
    (AB or CD) components, e.g. 'AB' -> ABx, ABy, ABz."""
    ops, outs = hrr_program(l1, l2)
    name = {}
    for (a, b) in [(a, (0, 0, 0)) for a in inp]:
        name[(a, b)] = inp[a]
    for n, (k, s0, s1, i) in enumerate(ops):
        v = f"{prefix}{n}"
        E(f"const {_R} {v} = {name[s0]} + {dvec}{AX[i]} * {name[s1]};")
        name[k] = v
    return [name[k] for k in outs]


def emit_sph_2(E, l1, l2, pure1, pure2, vals, prefix):
    """vals: Cartesian (a-major, b-minor) expressions; apply solid-harmonic
    transforms (standard order) on the pure indices.  Returns (n1, n2, list)."""
    n1c, n2c = ncart(l1), ncart(l2)
    cur = vals
    d1, d2 = n1c, n2c
    cnt = [0]

    def newv(expr):
        v = f"{prefix}{cnt[0]}"
        cnt[0] += 1
        E(f"const {_R} {v} = {expr};")
        return v

    if pure1 and l1 > 0:
        rows = sh_rows(l1)
        nxt = []
        for r in rows:
            for j in range(d2):
                terms = [f"{fmt(c)} * {cur[ic * d2 + j]}" for ic, c in r]
                nxt.append(newv(' + '.join(terms)))
        cur = nxt
        d1 = 2 * l1 + 1
    if pure2 and l2 > 0:
        rows = sh_rows(l2)
        nxt = []
        for i in range(d1):
            for r in rows:
                terms = [f"{fmt(c)} * {cur[i * d2 + ic]}" for ic, c in r]
                nxt.append(newv(' + '.join(terms)))
        cur = nxt
        d2 = 2 * l2 + 1
    return d1, d2, cur


def pure_variants(l1, l2):
    """Purity combinations that change the code: purity of an s shell is
    irrelevant."""
    v1 = [0, 1] if l1 > 0 else [0]
    v2 = [0, 1] if l2 > 0 else [0]
    return [(p1, p2) for p1 in v1 for p2 in v2]


# ---------------------------------------------------------------------------
# kernel emission
# ---------------------------------------------------------------------------


def _emit_stage2(E, sfx, gb, gk, count, act=None):
    """Screening, pfac, T and the Boys-table fetch of one primitive quartet
    (pair data in Pg{sfx}.. Qg{sfx}.., records gb / gk of the pair arrays).
    Screened primitives get pfac = 0: the recursion then produces exact
    zeros, which keeps the loop free of divergent branches; so do primitive
    quartets past the lane's own range (act false) in the wave-uniform loop."""
    E(f"const double gp{sfx} = Pg{sfx}.w, gq{sfx} = Qg{sfx}.w;")
    E(f"const double rs{sfx} = rsqrt(gp{sfx} + gq{sfx});")
    E(f"const double oos{sfx} = rs{sfx} * rs{sfx};")
    E(f"const double pf{sfx} = PAk{sfx}.w * QCk{sfx}.w * rs{sfx} * A.scale;")
    E(f"bool ok{sfx} = {act + ' && ' if act else ''}xb{sfx}.x + xk{sfx}.x > lnprec;")
    E.block("if (A.method == L2G_SCREEN_ORIGINAL)")
    E(f"ok{sfx} = ok{sfx} && !(fabs(pf{sfx}) < qprec);")
    E.end()
    E.block("else if (A.method == L2G_SCREEN_CONSERVATIVE)")
    E(f"const double nsb = A.pp_nonsph[{gb}], nsk = A.pp_nonsph[{gk}];")
    E(f"ok{sfx} = ok{sfx} && !(fabs(pf{sfx}) * fmax(1.0, nsb * nsk) * npbk < qprec);")
    E.end()
    if count is not None:
        E(f"if ({count}ok{sfx}) ++nsurv;")
    E(f"const double PQx{sfx} = Pg{sfx}.x - Qg{sfx}.x, PQy{sfx} = Pg{sfx}.y - Qg{sfx}.y, "
      f"PQz{sfx} = Pg{sfx}.z - Qg{sfx}.z;")
    if _ATT:
        # erf(w r)/r: G_m(T) = s^(m+1/2) F_m(sT); the screening keeps the
        # Coulomb prefactor, as the generic kernel and the reference do
        E(f"const double rho{sfx} = gp{sfx} * gq{sfx} * oos{sfx};")
        E(f"sw{sfx} = A.omega * A.omega / (A.omega * A.omega + rho{sfx});")
        E(f"pfac{sfx} = ok{sfx} ? pf{sfx} * sqrt(sw{sfx}) : 0.0;")
        E(f"boys_fetch(sw{sfx} * rho{sfx} * (PQx{sfx} * PQx{sfx} + PQy{sfx} * PQy{sfx} + "
          f"PQz{sfx} * PQz{sfx}), boysL, bq{sfx});")
    else:
        E(f"pfac{sfx} = ok{sfx} ? pf{sfx} : 0.0;")
        E(f"boys_fetch(gp{sfx} * gq{sfx} * oos{sfx} * (PQx{sfx} * PQx{sfx} + PQy{sfx} * PQy{sfx} + "
          f"PQz{sfx} * PQz{sfx}), boysL, bq{sfx});")


def emit_prim_loop(E, plan, graph, gi, accs):
    """Software-pipelined loop over the primitive quartets of one quartet:
    pair data of the next primitive are loaded at the top of an iteration and
    its Boys-table row halfway through the recursion, so both memory
    latencies overlap the FP64 work of the current primitive."""
    if _MFMA:
        # wave-uniform trip count (tmax, the wave's longest range): a lane
        # past its own range, or without primitive quartets, reads record 0
        # of the pair arrays (present whenever any lane of the wave has work)
        # and runs with pfac = 0, so every MFMA sees all 64 lanes active
        E.block("if (tmax > 0)")
        E("int ib = pbeg / npk1, ik = pbeg - ib * npk1;")
        E("const int gb = trip > 0 ? pb0 + ib : 0, gk = trip > 0 ? pk0 + ik : 0;")
        ga, ka = "gb", "gk"
    else:
        E.block("if (pend > pbeg)")
        E("int ib = pbeg / npk, ik = pbeg - ib * npk;")
        ga, ka = "pb0 + ib", "pk0 + ik"
    E(f"double4 Pg = A.pp_P[{ga}], PAk = A.pp_PA[{ga}];")
    E(f"double2 xb = A.pp_aux[{ga}];")
    E(f"double4 Qg = A.pp_P[{ka}], QCk = A.pp_PA[{ka}];")
    E(f"double2 xk = A.pp_aux[{ka}];")
    E("double pfac;")
    if _ATT:
        E("double sw;")
    E("L2gBoysArg bq;")
    E.block("")
    if _MFMA:
        _emit_stage2(E, "", ga, ka, "" if gi == 0 else None, act="trip > 0")
    else:
        _emit_stage2(E, "", ga, ka, "" if gi == 0 else None)
    E.end()
    if _MFMA:
        E.block("for (int t = 0; t < tmax; ++t)")
        E("int ikn = ik + 1, ibn = ib;")
        E("if (ikn >= npk) { ikn = 0; ++ibn; }")
        E("if (ibn >= npb) ibn = 0;")
        E("const bool actn = t + 1 < trip;")
        E("const int gbn = actn ? pb0 + ibn : 0, gkn = actn ? pk0 + ikn : 0;")
        gan, kan = "gbn", "gkn"
    else:
        E.block("for (int p = pbeg; p < pend; ++p)")
        E("int ikn = ik + 1, ibn = ib;")
        E("if (ikn == npk) { ikn = 0; ++ibn; }")
        E("if (ibn == npb) ibn = 0;  // past the end: any valid pair, result unused")
        gan, kan = "pb0 + ibn", "pk0 + ikn"
    E(f"const double4 Pgn = A.pp_P[{gan}], PAkn = A.pp_PA[{gan}];")
    E(f"const double2 xbn = A.pp_aux[{gan}];")
    E(f"const double4 Qgn = A.pp_P[{kan}], QCkn = A.pp_PA[{kan}];")
    E(f"const double2 xkn = A.pp_aux[{kan}];")
    E("double pfacn;")
    if _ATT:
        E("double swn;")
    E("L2gBoysArg bqn;")
    E(f"double F[{plan.L + 1}];")
    E(f"boys_eval<{plan.L}>(bq, pfac, F);")
    if _ATT and plan.L > 0:
        E.block("")
        E("double sm = sw;")
        E.block(f"for (int m = 1; m <= {plan.L}; ++m)")
        E("F[m] *= sm;")
        E("sm *= sw;")
        E.end()
        E.end()
    if _R == 'double':
        E("const double* Fv = F;")
    else:
        E(f"{_R} Fv[{plan.L + 1}];")
        for m in range(plan.L + 1):
            E(f"Fv[{m}] = ({_R})F[{m}];")
    if plan.L > 0:
        E("const double gp = Pg.w, gq = Qg.w;")
        E("const double rs = rsqrt(gp + gq);")
        E("const double oos = rs * rs;")
        E("const double PQx = Pg.x - Qg.x, PQy = Pg.y - Qg.y, PQz = Pg.z - Qg.z;")
        R = _R
        E(f"const {R} PAx = ({R})PAk.x, PAy = ({R})PAk.y, PAz = ({R})PAk.z;")
        E(f"const {R} QCx = ({R})QCk.x, QCy = ({R})QCk.y, QCz = ({R})QCk.z;")
        E("const double goq = gq * oos, gop = gp * oos;")
        E(f"const {R} WPx = ({R})(-goq * PQx), WPy = ({R})(-goq * PQy), WPz = ({R})(-goq * PQz);")
        E(f"const {R} WQx = ({R})(gop * PQx), WQy = ({R})(gop * PQy), WQz = ({R})(gop * PQz);")
        E(f"const {R} oo2z = ({R})xb.y, oo2e = ({R})xk.y, oo2ze = ({R})(0.5 * oos);")
        E(f"const {R} roz = ({R})goq, roe = ({R})gop;")
        E("(void)PAx; (void)PAy; (void)PAz; (void)QCx; (void)QCy; (void)QCz;")
        E("(void)WPx; (void)WPy; (void)WPz; (void)WQx; (void)WQy; (void)WQz;")
        E("(void)oo2z; (void)oo2e; (void)oo2ze; (void)roz; (void)roe;")
        emit_vrr_coefs(E, vrr_coef_names(plan, graph, gi))

    def mid():
        E.block("")
        if _MFMA:
            _emit_stage2(E, "n", gan, kan, "" if gi == 0 else None, act="actn")
        else:
            _emit_stage2(E, "n", gan, kan, "p + 1 < pend && " if gi == 0 else None)
        E.end()

    emit_vrr_group(E, plan, graph, gi, accs, mid=mid)
    E("Pg = Pgn; PAk = PAkn; xb = xbn; Qg = Qgn; QCk = QCkn; xk = xkn;")
    E("pfac = pfacn; bq = bqn; ib = ibn; ik = ikn;")
    if _ATT:
        E("sw = swn;")
    E.end()  # p loop
    E.end()  # npq > 0


def emit_class(plan: ClassPlan, block=128, mode='eri', waves=0, real='double', att=False, mfma=False):
    la, lb, lc, ld = plan.la, plan.lb, plan.lc, plan.ld
    graph = VrrGraph()
    ne, nf = len(plan.ecomps), len(plan.fcomps)
    G = len(plan.groups)
    E = Emitter()
    E(f"// class ({la}{lb}|{lc}{ld}): L={plan.L}, {ne}x{nf} contracted [e0|f0] targets, "
      f"{G} target group(s)")
    E(f"// group sizes (bra components x ket targets): {[(len(g), len(f)) for g, f in plan.groups]}")
    E(f"// group graph nodes: {plan.group_nodes}; estimated peak live doubles: {plan.group_peak}")
    global _R, _ATT, _MFMA
    _R = real
    _ATT = att
    _MFMA = mfma and real == 'double'
    if _MFMA:
        E("// contraction on the matrix cores: every target accumulates its primitive")
        E("// quartets through v_mfma_f64_4x4x4_4b_f64 with A = the lane's value, B a")
        E("// one-hot selector that returns each lane's value to its own D slot, and")
        E("// C/D the accumulator (AGPRs, no VALU moves); loops are wave-uniform")
    kname = f"l2g_{'erf' if att else mode}{'32' if real == 'float' else ''}_{plan.name}"
    attr = f" __attribute__((amdgpu_waves_per_eu({waves}, {waves})))" if waves else ""
    if mode == 'fock':
        E("// Fock variant: digests each integral of the unique quartet into the")
        E("// unsymmetrised Coulomb (jg) and exchange (kg) accumulators with the")
        E("// degeneracy factor, as compute_2body_fock does")
        E("// (tests/hartree-fock/hartree-fock++.cc:1709-1743); J = (jg + jg^T)/4,")
        E("// K = (kg + kg^T)/8, G = 2J - K after symmetrisation.")
    if mode == 'schwarz':
        E("// Schwarz-bound variant: work item = shell pair p, evaluates (p|p) and")
        E("// stores sqrt(max|.|) (norm 0) or sqrt(||.||_2) (norm 1) at out[out_off]")
        E("// (tests/hartree-fock/hartree-fock++.cc:1231-1298).")
    E(f"__global__ void __launch_bounds__({block}){attr} {kname}(const L2gEriArgs A)")
    E("{")
    E.ind += 1
    E("const int tid = blockIdx.x * blockDim.x + threadIdx.x;")
    E("const int nthr = gridDim.x * blockDim.x;")
    if _MFMA:
        E("const int lane = threadIdx.x & 63;")
        E(f"const double msel = {MFMA_SELECTOR};")
        E(f"const int msrc = {MFMA_SOURCE};")
        if mode in ('schwarz', 'eri'):
            E("// A.coop: a wave per work item, each lane taking a contiguous 1/64 of the")
            E("// primitive quartets; the group accumulators are summed over the wave")
            E("const bool coop = A.coop != 0;")
            E("// the wave walks the work list together: lanes past its end take the last")
            E("// item as a stand-in (all work, no stores), so every loop stays uniform")
            E.block("for (int qb = coop ? (tid >> 6) : tid - lane; qb < A.n; qb += coop ? (nthr >> 6) : nthr)")
            E("const bool live = coop || qb + lane < A.n;")
            E("const int qi = coop ? qb : (live ? qb + lane : A.n - 1);")
        else:
            E("// the wave walks the work list together: lanes past its end take the last")
            E("// item as a stand-in (all work, no digest), so every loop stays uniform")
            E.block("for (int qb = tid - lane; qb < A.n; qb += nthr)")
            E("const bool live = qb + lane < A.n;")
            E("const int qi = live ? qb + lane : A.n - 1;")
    elif mode in ('schwarz', 'eri'):
        E("// A.coop: a wave per work item, each lane taking a contiguous 1/64 of the")
        E("// primitive quartets; the group accumulators are summed over the wave")
        E("const bool coop = A.coop != 0;")
        E("const int lane = threadIdx.x & 63;")
        E.block("for (int qi = coop ? (tid >> 6) : tid; qi < A.n; qi += coop ? (nthr >> 6) : nthr)")
    else:
        E.block("for (int qi = tid; qi < A.n; qi += nthr)")
    E("const int bp = A.bra_pair[qi];")
    E("const int kp = " + ("bp;" if mode == 'schwarz' else "A.ket_pair[qi];"))
    E("const int meta = A.meta[qi];")
    E("const int pb0 = A.pair_first[bp], npb = A.pair_np[bp];")
    E("const int pk0 = A.pair_first[kp], npk = A.pair_np[kp];")
    E("const double ABx = A.pair_AB[3 * bp], ABy = A.pair_AB[3 * bp + 1], ABz = A.pair_AB[3 * bp + 2];")
    E("const double CDx = A.pair_AB[3 * kp], CDy = A.pair_AB[3 * kp + 1], CDz = A.pair_AB[3 * kp + 2];")
    E(f"const double* __restrict__ boysL = A.boys + (size_t){plan.L} * L2G_BOYS_NINT * 8;")
    E("const double npbk = (double)npb * (double)npk;")
    E("// primitive quartets [pbeg, pend) of this work item (p = ib * npk + ik)")
    E("int pbeg = 0, pend = npb * npk;")
    if mode == 'fock':
        E.block("if (A.q_prange)")
        E("const int2 pr = A.q_prange[qi];")
        E("pbeg = pr.x;")
        E("pend = pr.y;")
        E.end()
    if mode in ('schwarz', 'eri'):
        E.block("if (coop)")
        E("const int len = pend - pbeg, b0 = pbeg + (int)(((long long)len * lane) >> 6);")
        E("pend = pbeg + (int)(((long long)len * (lane + 1)) >> 6);")
        E("pbeg = b0;")
        E.end()
    if _MFMA:
        E("// wave-uniform primitive loop: the longest range of the wave")
        E("const int trip = pend - pbeg;")
        E("int tmax = trip;")
        E("for (int o = 32; o > 0; o >>= 1) tmax = max(tmax, __shfl_xor(tmax, o, 64));")
        E("const int npk1 = npk > 0 ? npk : 1;")
    if mode == 'fock':
        E("const double lnprec = A.q_lnprec[qi];")
        E("const double qprec = A.method == L2G_SCREEN_ORIGINAL || A.method == L2G_SCREEN_CONSERVATIVE ? exp(lnprec) : 0.0;")
    else:
        E("const double lnprec = A.ln_precision;")
        E("const double qprec = A.precision;")
    E("(void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz; (void)npbk;")
    E("// lane-private scratch column; the empty asm keeps the compiler from")
    E("// hoisting one address per row/column out of the quartet loop")
    E("double* yb = A.ybuf + tid;")
    E("asm volatile(\"\" : \"+v\"(yb));")
    E("(void)yb;")

    # storage of ket-transformed rows: registers when the whole target fits
    # one group and the row block is small, else a lane-private HBM column
    ncol_max = ncart(lc) * ncart(ld)
    in_regs = (G == 1) and ne * ncol_max <= 48 and not plan.split_rows
    if in_regs:
        E(f"{_R} Y[{ne}][{ncol_max}];")
    if _R != 'double':
        E(f"const {_R} fABx = ({_R})ABx, fABy = ({_R})ABy, fABz = ({_R})ABz;")
        E(f"const {_R} fCDx = ({_R})CDx, fCDy = ({_R})CDy, fCDz = ({_R})CDz;")
    E("int nsurv = 0;  // primitive quartets that pass screening")
    if mode == 'schwarz':
        E("double hmax = 0.0;  // max |[e0|e0]| over the pre-HRR diagonal")
    xoff = ne * ncol_max  # raw-row region of the scratch column (split rows)
    # ERI output staged in the scratch column after the row regions
    staged = mode == 'eri' and not in_regs
    stage_off = xoff + (ne * nf if plan.split_rows else 0)
    for gi, (grp, gfs) in enumerate(plan.groups):
        ngf = len(gfs)
        accs = [f"acc{gi}_{t}" for t in range(len(grp) * ngf)]
        E(f"// ---- target group {gi}: bra components {[plan.ecomps[i] for i in grp]}, "
          f"{ngf} of {nf} ket targets")
        E.block("")
        for a in accs:
            E(f"{_R} {a} = 0;")
        emit_prim_loop(E, plan, graph, gi, accs)
        if _MFMA:
            # the contracted values back to their own lanes (a wave sum is
            # the same over any lane order)
            if mode in ('schwarz', 'eri'):
                E.block("if (!coop)")
            else:
                E.block("")
            for a in accs:
                E(f"{a} = __shfl({a}, msrc, 64);")
            E.end()
        if mode in ('schwarz', 'eri'):
            E.block("if (coop)")
            for a in accs:
                E(f"{a} = l2g_wave_sum({a});")
            E.end()
        if mode == 'schwarz':
            for r, ie in enumerate(grp):
                for t, jf in enumerate(gfs):
                    if plan.ecomps[ie] == plan.fcomps[jf]:
                        E(f"hmax = fmax(hmax, fabs((double){accs[r * ngf + t]}));")
        # ket HRR + ket spherical transform for each row of the group; a row
        # cut across groups is parked raw and transformed after its last piece
        full = plan.full_rows(gi)
        if not full:
            for r, ie in enumerate(grp):
                for t, jf in enumerate(gfs):
                    E(f"yb[(size_t)({xoff + ie * nf + jf}) * A.ystride] = {accs[r * ngf + t]};")
        for r, ie in enumerate(grp):
            if not full and plan.row_last_group(ie) != gi:
                continue
            E.block("")
            if full:
                inp = {plan.fcomps[jf]: accs[r * ngf + t] for t, jf in enumerate(gfs)}
            else:
                for jf in range(nf):
                    E(f"const {_R} xr{jf} = ({_R})yb[(size_t)({xoff + ie * nf + jf}) * A.ystride];")
                inp = {plan.fcomps[jf]: f"xr{jf}" for jf in range(nf)}
            cvals = emit_hrr(E, lc, ld, inp, "h", "CD" if _R == 'double' else "fCD")
            E.block("switch ((meta >> 5) & 3)")
            for (pc, pd) in pure_variants(lc, ld):
                E.block(f"case {pc | (pd << 1)}:")
                d1, d2, vals = emit_sph_2(E, lc, ld, pc, pd, cvals, "k")
                for col, v in enumerate(vals):
                    if in_regs:
                        E(f"Y[{ie}][{col}] = {v};")
                    else:
                        E(f"yb[(size_t)({ie * ncol_max + col}) * A.ystride] = {v};")
                E("break;")
                E.end()
            E("default: break;")
            E.end()
            E.end()
        E.end()  # group block
    # bra HRR + bra sph + permuted store, column by column over the ket
    if _MFMA:
        E("if (!live) continue;  // a stand-in lane: no store, no digest")
    if mode == 'eri':
        E("if (coop) nsurv = l2g_wave_sum(nsurv);")
        E("if (A.nz && (!coop || lane == 0)) A.nz[A.qidx[qi]] = nsurv > 0 ? 1 : 0;")
    elif mode == 'schwarz':
        E("double smax = 0.0, ssq = 0.0;")
    else:
        E("if (nsurv == 0) continue;  // every primitive screened: libint returns nullptr, nothing to digest")
        E("const int4 bf = A.q_bf[qi];")
        E("const double deg = A.q_scale[qi];")
        E("// matrices of this work item's request: D, jg, kg share one offset")
        E("const L2gMatDesc md = A.mats[A.q_req ? A.q_req[qi] : 0];")
        E("const long long N = md.n;")
        E("const double* __restrict__ Dm = A.dens + md.off;")
        E("double* __restrict__ jgm = A.jg + md.off;")
        E("double* __restrict__ kgm = A.kg + md.off;")
    E("// ---- bra HRR, bra spherical transform and permuted store")
    E(f"const L2gStore st = make_store(meta, A, qi, {la}, {lb}, {lc}, {ld});")
    E("const int ncol = st.nk1 * st.nk2;")

    def bra_column(colexpr, ysrc):
        E("const int ic = (" + colexpr + ") / st.nk2, id = (" + colexpr + ") - ic * st.nk2;")
        E(f"const size_t colofs = st.base + (size_t)st.pos3(ic) * st.s3 + (size_t)st.pos4(id) * st.s4;")
        E("const double colscale = st.uniform ? st.scale3(ic) * st.scale4(id) : 1.0;")
        if mode == 'fock':
            E("const long long bk3 = bf.z + st.pos3(ic), bk4 = bf.w + st.pos4(id);")
            E("const double d34 = Dm[bk3 * N + bk4];")
        inp = {e: ysrc(ie) for ie, e in enumerate(plan.ecomps)}
        bvals = emit_hrr(E, la, lb, inp, "g", "AB" if _R == 'double' else "fAB")
        E.block("switch ((meta >> 3) & 3)")
        for (pa, pb) in pure_variants(la, lb):
            E.block(f"case {pa | (pb << 1)}:")
            d1, d2, vals = emit_sph_2(E, la, lb, pa, pb, bvals, "b")
            if mode == 'fock':
                for i1 in range(d1):
                    E(f"const long long ba1_{i1} = bf.x + st.pos(st.l1, st.p1, {i1});")
                for i2 in range(d2):
                    E(f"const long long ba2_{i2} = bf.y + st.pos(st.l2, st.p2, {i2});")
                E("double jcol = 0.0;")
                for i1 in range(d1):
                    E(f"double k13_{i1} = 0.0, k14_{i1} = 0.0;")
                for i2 in range(d2):
                    E(f"double k24_{i2} = 0.0, k23_{i2} = 0.0;")
                for i2 in range(d2):
                    E(f"const double d24_{i2} = Dm[ba2_{i2} * N + bk4], d23_{i2} = Dm[ba2_{i2} * N + bk3];")
                for i1 in range(d1):
                    E(f"const double d13_{i1} = Dm[ba1_{i1} * N + bk3], d14_{i1} = Dm[ba1_{i1} * N + bk4];")
                for i1 in range(d1):
                    for i2 in range(d2):
                        v = vals[i1 * d2 + i2]
                        E(f"jcol += d12_{i1}_{i2} * {v}; j12_{i1}_{i2} += d34 * {v};")
                        E(f"k13_{i1} += d24_{i2} * {v}; k14_{i1} += d23_{i2} * {v};")
                        E(f"k24_{i2} += d13_{i1} * {v}; k23_{i2} += d14_{i1} * {v};")
                E("l2g_acc(&jgm[bk3 * N + bk4], deg * jcol, A.fx_hoff);")
                for i1 in range(d1):
                    E(f"l2g_acc(&kgm[ba1_{i1} * N + bk3], deg * k13_{i1}, A.fx_hoff);")
                    E(f"l2g_acc(&kgm[ba1_{i1} * N + bk4], deg * k14_{i1}, A.fx_hoff);")
                for i2 in range(d2):
                    E(f"l2g_acc(&kgm[ba2_{i2} * N + bk4], deg * k24_{i2}, A.fx_hoff);")
                    E(f"l2g_acc(&kgm[ba2_{i2} * N + bk3], deg * k23_{i2}, A.fx_hoff);")
            else:
                for i1 in range(d1):
                    for i2 in range(d2):
                        v = vals[i1 * d2 + i2]
                        if mode == 'eri' and staged:
                            # a wave-cooperative item spreads its columns over the lanes
                            E(f"if (coop) st.put(A.out, colofs, {i1}, {i2}, colscale, {v}); "
                              f"else st.put_staged(yb, A.ystride, {stage_off}, colofs, {i1}, {i2}, colscale, {v});")
                        elif mode == 'eri':
                            E(f"st.put(A.out, colofs, {i1}, {i2}, colscale, {v});")
                        else:
                            E(f"smax = fmax(smax, fabs({v})); ssq += {v} * {v};")
            E("break;")
            E.end()
        E("default: break;")
        E.end()

    if mode == 'fock':
        # D(ab) block and the J(ab) accumulator of the bra, sized for the
        # Cartesian case (the pure case uses the leading part)
        n1m, n2m = ncart(la), ncart(lb)
        for i1 in range(n1m):
            for i2 in range(n2m):
                E(f"double j12_{i1}_{i2} = 0.0;")
        E.block("switch ((meta >> 3) & 3)")
        for (pa, pb) in pure_variants(la, lb):
            d1 = 2 * la + 1 if pa else ncart(la)
            d2 = 2 * lb + 1 if pb else ncart(lb)
            E(f"case {pa | (pb << 1)}: break;")
        E("default: break;")
        E.end()
        for i1 in range(n1m):
            for i2 in range(n2m):
                E(f"const double d12_{i1}_{i2} = ({i1} < st.n1 && {i2} < st.n2) ? "
                  f"Dm[(bf.x + st.pos(st.l1, st.p1, {i1})) * N + bf.y + st.pos(st.l2, st.p2, {i2})] : 0.0;")

    if in_regs:
        for col in range(ncol_max):
            if mode == 'eri':
                E.block(f"if ({col} < ncol && (!coop || lane == {col % 64}))")
            else:
                E.block(f"if ({col} < ncol)")
            bra_column(str(col), lambda ie, c=col: f"Y[{ie}][{c}]")
            E.end()
    else:
        if mode == 'eri':  # a wave-cooperative item spreads its columns over the lanes
            E.block("for (int col = coop ? lane : 0; col < ncol; col += coop ? 64 : 1)")
        else:
            E.block("for (int col = 0; col < ncol; ++col)")
        bra_column("col", lambda ie: f"({_R})yb[((size_t)({ie * ncol_max}) + col) * A.ystride]")
        E.end()
    if mode == 'eri' and staged:
        E(f"if (!coop) st.flush(A.out, yb, A.ystride, {stage_off});")
    if mode == 'schwarz':
        E.block("if (!coop || lane == 0)")
        E("A.out[A.out_off[qi]] = A.schwarz_norm ? sqrt(sqrt(ssq)) : sqrt(smax);")
        E("if (A.out2) A.out2[A.out_off[qi]] = sqrt(hmax);")
        E.end()
    if mode == 'fock':
        for i1 in range(ncart(la)):
            for i2 in range(ncart(lb)):
                E(f"if ({i1} < st.n1 && {i2} < st.n2) l2g_acc(&jgm[(bf.x + st.pos(st.l1, st.p1, {i1})) * N + "
                  f"bf.y + st.pos(st.l2, st.p2, {i2})], deg * j12_{i1}_{i2}, A.fx_hoff);")
    E.end()  # quartet loop
    E.ind -= 1
    E("}")
    _R = 'double'
    _ATT = False
    _MFMA = False
    return kname, E.text(), in_regs


def canonical_classes(lmax):
    out = []
    pairs = [(a, b) for a in range(lmax + 1) for b in range(a + 1)]
    for (la, lb) in pairs:
        for (lc, ld) in pairs:
            if la + lb <= lc + ld:
                out.append((la, lb, lc, ld))
    return out


def class_waves(plan, budget, waves):
    """waves_per_eu hint: explicit, or 0 (compiler default)."""
    return waves


def generate(outdir, lmax=2, budget=160, nfiles=8, waves=0, only=None, budgets=None,
             modes=('eri', 'fock', 'schwarz', 'fp32', 'erf'), generic_lmax=3, max_nodes=0, fp32_lmax=2,
             erf_lmax=2, mfma_acc=False):
    """Write the kernels and the dispatch tables.  `budgets` maps a class name
    ('2222') to its own (budget, waves).  Classes with an l above 2 whose
    partitioned recursion evaluates more than `max_nodes` nodes per primitive
    quartet (0: no limit) get no kernel of their own and run on the generic
    kernel; FP32 variants are made for classes with every l <= fp32_lmax, and
    erf-attenuated ERI variants (mode 'erf') for every l <= erf_lmax."""
    os.makedirs(outdir, exist_ok=True)
    classes = canonical_classes(lmax)
    if only:
        classes = [c for c in classes if "".join(map(str, c)) in only]
    budgets = budgets or {}

    def cfg(c):
        v = budgets.get("".join(map(str, c)), (budget, waves))
        return tuple(v) + (0,) * (4 - len(v))  # (budget, waves, FP32 pays, generic for few deep quartets)

    plans = [plan_class(*c, budget=cfg(c)[0]) for c in classes]
    if max_nodes:
        plans = [p for p in plans if max(p.la, p.lb, p.lc, p.ld) <= 2 or sum(p.group_nodes) <= max_nodes]
    # reduced-precision variants hold twice as many values per register
    plans32 = [plan_class(p.la, p.lb, p.lc, p.ld, budget=2 * cfg((p.la, p.lb, p.lc, p.ld))[0]) for p in plans
               if max(p.la, p.lb, p.lc, p.ld) <= fp32_lmax] if 'fp32' in modes else []
    jobs = []  # (plan, mode, real)
    for p in plans:
        if 'eri' in modes:
            jobs.append((p, 'eri', 'double'))
        if 'fock' in modes:
            jobs.append((p, 'fock', 'double'))
        if 'schwarz' in modes and (p.la, p.lb) == (p.lc, p.ld):
            jobs.append((p, 'schwarz', 'double'))
    for p in plans:
        if 'erf' in modes and max(p.la, p.lb, p.lc, p.ld) <= erf_lmax:
            jobs.append((p, 'erf', 'double'))
    for p in plans32:
        if 'eri' in modes:
            jobs.append((p, 'eri', 'float'))
        if 'fock' in modes:
            jobs.append((p, 'fock', 'float'))
    weights = sorted(((sum(p.group_nodes) * (0.6 if m == 'schwarz' else 1.0) + 200, i)
                      for i, (p, m, r) in enumerate(jobs)), reverse=True)
    files = [[] for _ in range(nfiles)]
    load = [0] * nfiles
    for w, i in weights:
        j = load.index(min(load))
        files[j].append(i)
        load[j] += w
    tables = {('eri', 'double'): [], ('schwarz', 'double'): [], ('fock', 'double'): [],
              ('eri', 'float'): [], ('fock', 'float'): [], ('erf', 'double'): []}
    written = []
    for fi, idxs in enumerate(files):
        fn = os.path.join(outdir, f"eri_gen_{fi}.hip")
        parts = ['// generated by libint_amd/codegen/gen_eri.py — do not edit',
                 '#include "eri_device.h"', '']
        for i in sorted(idxs):
            p, mode, real = jobs[i]
            wv = cfg((p.la, p.lb, p.lc, p.ld))[1]
            kname, code, in_regs = emit_class(p, mode='eri' if mode == 'erf' else mode, waves=wv, real=real,
                                              att=mode == 'erf', mfma=mfma_acc)
            parts.append(code)
            parts.append(f"void {kname}_launch(const L2gEriArgs& A, int grid, int block, hipStream_t s) {{")
            parts.append(f"  hipLaunchKernelGGL({kname}, dim3(grid), dim3(block), 0, s, A);")
            parts.append("}")
            parts.append("")
            tables[(mode, real)].append((p, kname, in_regs))
        new = '\n'.join(parts)
        if not (os.path.exists(fn) and open(fn).read() == new):
            with open(fn, 'w') as f:
                f.write(new)
        written.append(fn)
    table, stable, ftable = tables[('eri', 'double')], tables[('schwarz', 'double')], tables[('fock', 'double')]
    table32, ftable32 = tables[('eri', 'float')], tables[('fock', 'float')]
    etable = tables[('erf', 'double')]
    lines = ['// generated by libint_amd/codegen/gen_eri.py — do not edit',
             '#include "eri_device.h"', '']
    for p, k, _ in table + stable + ftable + table32 + ftable32 + etable:
        lines.append(f"void {k}_launch(const L2gEriArgs& A, int grid, int block, hipStream_t s);")
    lines.append('')
    lines.append(f"extern const int l2g_gen_lmax = {lmax};")
    for name, tab in (("l2g_gen_classes", table), ("l2g_gen_schwarz", stable), ("l2g_gen_fock", ftable),
                      ("l2g_gen_classes32", table32), ("l2g_gen_fock32", ftable32),
                      ("l2g_gen_classes_erf", etable)):
        lines.append(f"extern const L2gClassInfo {name}[] = {{")
        for p, k, in_regs in tab:
            ybuf = 0 if in_regs else len(p.ecomps) * ncart(p.lc) * ncart(p.ld) + \
                (len(p.ecomps) * len(p.fcomps) if p.split_rows else 0)
            if not in_regs and name in ("l2g_gen_classes", "l2g_gen_classes32", "l2g_gen_classes_erf"):
                ybuf += ncart(p.la) * ncart(p.lb) * ncart(p.lc) * ncart(p.ld)  # staged output
            pays = cfg((p.la, p.lb, p.lc, p.ld))[2] if name in ("l2g_gen_classes32", "l2g_gen_fock32") else 0
            fewgen = cfg((p.la, p.lb, p.lc, p.ld))[3] if name == "l2g_gen_classes" else 0
            lines.append(f"  {{{p.la}, {p.lb}, {p.lc}, {p.ld}, {len(p.groups)}, {ybuf}, "
                         f"{sum(p.group_nodes)}, {pays}, {fewgen}, &{k}_launch}},")
        if not tab:
            lines.append("  {0, 0, 0, 0, 0, 0, 0, 0, 0, nullptr},")
        lines.append("};")
    lines.append(f"extern const int l2g_gen_nclasses = {len(table)};")
    lines.append(f"extern const int l2g_gen_nschwarz = {len(stable)};")
    lines.append(f"extern const int l2g_gen_nfock = {len(ftable)};")
    lines.append(f"extern const int l2g_gen_nclasses32 = {len(table32)};")
    lines.append(f"extern const int l2g_gen_nfock32 = {len(ftable32)};")
    lines.append(f"extern const int l2g_gen_nclasses_erf = {len(etable)};")
    fn = os.path.join(outdir, "eri_table.cpp")
    new = '\n'.join(lines) + '\n'
    if not (os.path.exists(fn) and open(fn).read() == new):
        with open(fn, 'w') as f:
            f.write(new)
    written.append(fn)
    # programs of the generic kernel: every canonical class up to generic_lmax
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import gen_generic
    fn = os.path.join(outdir, "generic_table.cpp")
    new = gen_generic.emit_tables(canonical_classes(generic_lmax), generic_lmax)
    if not (os.path.exists(fn) and open(fn).read() == new):
        with open(fn, 'w') as f:
            f.write(new)
    written.append(fn)
    return written, plans


if __name__ == '__main__':
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("outdir")
    ap.add_argument("--lmax", type=int, default=2)
    ap.add_argument("--budget", type=int, default=160)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--only", default="", help="comma-separated class names, e.g. 2222")
    ap.add_argument("--class-cfg", default="",
                    help="per class budget:waves[:fp32], e.g. 2222=64:2,2122=80:2:1 (fp32 1: the class's FP32 "
                         "variant measured faster than its FP64 kernel, so mixed batches route quartets to it)")
    ap.add_argument("--nfiles", type=int, default=8)
    ap.add_argument("--modes", default="eri,fock,schwarz,fp32,erf")
    ap.add_argument("--generic-lmax", type=int, default=3, help="highest l of the generic kernel's classes")
    ap.add_argument("--max-nodes", type=int, default=0,
                    help="classes with l > 2 and more recursion nodes than this run on the generic kernel")
    ap.add_argument("--fp32-lmax", type=int, default=2, help="highest l with FP32 kernel variants")
    ap.add_argument("--mfma-acc", type=int, default=0,
                    help="1: FP64 kernels accumulate the contraction on the matrix cores")
    a = ap.parse_args()
    per = {}
    for item in filter(None, a.class_cfg.split(",")):
        k, v = item.split("=")
        per[k] = tuple(int(x) for x in v.split(":"))
    files, plans = generate(a.outdir, a.lmax, a.budget, a.nfiles, a.waves,
                            set(filter(None, a.only.split(","))) or None, per, tuple(a.modes.split(",")),
                            a.generic_lmax, a.max_nodes, a.fp32_lmax, mfma_acc=bool(a.mfma_acc))
    for p in plans:
        print(p.name, 'groups', len(p.groups), 'nodes', sum(p.group_nodes), 'peak', max(p.group_peak))
