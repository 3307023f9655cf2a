// Generic (table-driven) kernel for the classes the register-resident
// kernels do not cover: every class with an f shell.  One 256-thread block
// evaluates one contracted shell quartet at a time with all intermediates in
// LDS; the per-class work comes from the programs of codegen/gen_generic.py:
//
//   1. per primitive quartet: screening and prefactors (every thread, the
//      values are block-uniform), F_m(T) by Chebyshev interpolation of each
//      m on its own thread (the scheme of FmEval_Chebyshev7::eval,
//      include/libint2/boys.h:345-454), then the OS VRR level by level, a
//      thread computing whole m-vectors of [e0|f0] nodes and adding the
//      m = 0 value of target nodes to their contracted accumulator;
//   2. ket HRR (CD) and ket solid harmonics, one thread per bra row;
//   3. bra HRR (AB) and bra solid harmonics, one thread per ket column,
//      followed by the caller-order store (ERI), the J/K digest into LDS
//      accumulators flushed to the J/K arena (Fock; FP64 atomics, or
//      fixed-point words for the deterministic digest), or the norm
//      reduction (Schwarz).
//
// The maths is that of the generated kernels (engine.impl.h:1306-1985 for
// the reference's steps); only the mapping onto the GPU differs.
#include "l2g_internal.h"

namespace {

constexpr int kModeEri = 0, kModeFock = 1, kModeSchwarz = 2;

// solid-harmonic coefficients, standard order (m = -l..l), l <= 3:
// [l][m + l][Cartesian index]
__constant__ double c_sh[4][7][10];

__device__ __forceinline__ int ncart_d(int l) { return (l + 1) * (l + 2) / 2; }

// Fock digest accumulators in LDS: FP64 atomics, or (fx) fixed-point pairs
// whose high words sit hoff words after the low words (l2g_fx_add)
__device__ __forceinline__ void lds_acc(double* p, double v, bool fx, int hoff) {
  if (!fx) {
    atomicAdd(p, v);
    return;
  }
  unsigned long long* lo = reinterpret_cast<unsigned long long*>(p);
  l2g_fx_add(lo + hoff, lo, v);
}

__device__ __forceinline__ double lds_value(const double* p, bool fx, int hoff) {
  if (!fx) return *p;
  const unsigned long long* lo = reinterpret_cast<const unsigned long long*>(p);
  return l2g_fx_value(lo[hoff], lo[0]);
}

// F_m(T) for one m: the Chebyshev-7 interpolant of the interval holding T,
// or the asymptotic form above T_crit (boys.h:348-358)
__device__ __forceinline__ double boys_one(double T, int m, const double* __restrict__ tab) {
  if (T > L2G_BOYS_TCRIT) {
    const double oox = 1.0 / T;
    double f = 0.88622692545275801365 * sqrt(oox);
    for (int j = 1; j <= m; ++j) f *= ((double)j - 0.5) * oox;
    return f;
  }
  const double x7 = T * 7.0;
  const double ivd = fmin(floor(x7), (double)(L2G_BOYS_NINT - 1));
  const double xd = x7 - ivd - 0.5;
  const double* c = tab + ((size_t)m * L2G_BOYS_NINT + (size_t)ivd) * 8;
  double f = c[7];
  for (int k = 6; k >= 0; --k) f = fma(f, xd, c[k]);
  return f;
}

// out[i][j] (n1o x n2) = sum_a U_l1[i][a] in[a][j] when pure, else a copy
__device__ void sph_first(const double* in, double* out, int l1, int n2, int pure) {
  const int nc = ncart_d(l1);
  if (!pure || l1 == 0) {
    for (int k = 0; k < nc * n2; ++k) out[k] = in[k];
    return;
  }
  for (int i = 0; i < 2 * l1 + 1; ++i)
    for (int j = 0; j < n2; ++j) {
      double s = 0.0;
      for (int a = 0; a < nc; ++a) s += c_sh[l1][i][a] * in[a * n2 + j];
      out[i * n2 + j] = s;
    }
}

// out[i][j] (n1 x n2o) = sum_b U_l2[j][b] in[i][b] when pure, else a copy
__device__ void sph_second(const double* in, double* out, int n1, int l2, int pure) {
  const int nc = ncart_d(l2);
  const int no = pure && l2 > 0 ? 2 * l2 + 1 : nc;
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < no; ++j) {
      double s;
      if (no == nc && !(pure && l2 > 0)) {
        s = in[i * nc + j];
      } else {
        s = 0.0;
        for (int b = 0; b < nc; ++b) s += c_sh[l2][j][b] * in[i * nc + b];
      }
      out[i * no + j] = s;
    }
}

// NT threads per block: 1024 or 512 for classes whose LDS leaves room for
// one or two blocks per CU, so that every SIMD still holds 4 waves to hide
// the LDS and global-memory latencies of the recursion; 256 otherwise.
template <int MODE, int NT>
__global__ void __launch_bounds__(NT) l2g_generic_kernel(const L2gEriArgs A, const L2gGenericClass C,
                                                           const int4* __restrict__ prog, const int* __restrict__ idx) {
  constexpr int kThreads = NT;
  extern __shared__ double lds[];
  __shared__ double red[2 * NT / 64];
  const int t = threadIdx.x;
  const int ne = C.ne, nf = C.nf;
  const int la = C.la, lb = C.lb, lc = C.lc, ld = C.ld;
  for (int qi = blockIdx.x; qi < A.n; qi += gridDim.x) {
    const int bp = A.bra_pair[qi];
    const int kp = MODE == kModeSchwarz ? bp : A.ket_pair[qi];
    const int meta = A.meta[qi];
    const int pb0 = A.pair_first[bp], npb = A.pair_np[bp];
    const int pk0 = A.pair_first[kp], npk = A.pair_np[kp];
    const double AB[3] = {A.pair_AB[3 * bp], A.pair_AB[3 * bp + 1], A.pair_AB[3 * bp + 2]};
    const double CD[3] = {A.pair_AB[3 * kp], A.pair_AB[3 * kp + 1], A.pair_AB[3 * kp + 2]};
    const double npbk = (double)npb * (double)npk;
    double lnprec, qprec;
    if (MODE == kModeFock) {
      lnprec = A.q_lnprec[qi];
      qprec = A.method == L2G_SCREEN_ORIGINAL || A.method == L2G_SCREEN_CONSERVATIVE ? exp(lnprec) : 0.0;
    } else {
      lnprec = A.ln_precision;
      qprec = A.precision;
    }
    int pbeg = 0, pend = npb * npk;
    const int gpart = MODE == kModeEri ? A.gpart_mode : 0;
    if ((MODE == kModeFock || gpart == 1) && A.q_prange) {
      const int2 pr = A.q_prange[qi];
      pbeg = pr.x;
      pend = pr.y;
    }
    int nsurv = 0;
    if (gpart == 2) {
      // the quartet's pieces, summed in piece order; no primitive loop here
      const int r0 = A.gpart_first[qi], r1 = A.gpart_first[qi + 1];
      const int nacc = ne * C.accs;
      for (int i = t; i < nacc; i += kThreads) {
        double s = 0.0;
        for (int r = r0; r < r1; ++r) s += A.gpart[(long long)r * A.gpart_stride + i];
        lds[C.acc + i] = s;
      }
      for (int r = r0; r < r1; ++r) nsurv += (int)A.gpart[(long long)r * A.gpart_stride + nacc];
      pend = pbeg;
    } else {
      for (int i = t; i < ne * C.accs; i += kThreads) lds[C.acc + i] = 0.0;
    }
    if (t < 16) {
      lds[C.zero + t] = 0.0;
      if (C.dual) lds[C.zero + C.pstride + t] = 0.0;  // the second block's zero vector
    }
    __syncthreads();
    // ---- 1. primitive quartets: prefactors, F_m, VRR, contraction.  With
    // C.dual the next two surviving primitive quartets go through the
    // recursion together, the second one's coefficients and level regions
    // C.pstride doubles after the first's; each node task computes both and
    // adds them to the accumulator in primitive order.
    const int slots = C.dual ? 2 : 1;
    int p = pbeg;
    while (p < pend) {
      int ns = 0;  // surviving primitive quartets gathered for this pass (block-uniform)
      for (; p < pend && ns < slots; ++p) {
        const int ib = p / npk, ik = p - ib * npk;
        const double4 Pg = A.pp_P[pb0 + ib], PA = A.pp_PA[pb0 + ib];
        const double2 xb = A.pp_aux[pb0 + ib];
        const double4 Qg = A.pp_P[pk0 + ik], QC = A.pp_PA[pk0 + ik];
        const double2 xk = A.pp_aux[pk0 + ik];
        const double gp = Pg.w, gq = Qg.w;
        const double rs = rsqrt(gp + gq);
        const double oos = rs * rs;
        const double pf = PA.w * QC.w * rs * A.scale;
        bool ok = xb.x + xk.x > lnprec;
        if (A.method == L2G_SCREEN_ORIGINAL) {
          ok = ok && !(fabs(pf) < qprec);
        } else if (A.method == L2G_SCREEN_CONSERVATIVE) {
          const double nsb = A.pp_nonsph[pb0 + ib], nsk = A.pp_nonsph[pk0 + ik];
          ok = ok && !(fabs(pf) * fmax(1.0, nsb * nsk) * npbk < qprec);
        }
        if (!ok) continue;  // block-uniform
        ++nsurv;
        const int so = ns * C.pstride;  // this primitive quartet's block
        ++ns;
        const double PQx = Pg.x - Qg.x, PQy = Pg.y - Qg.y, PQz = Pg.z - Qg.z;
        const double T = gp * gq * oos * (PQx * PQx + PQy * PQy + PQz * PQz);
        if (t < 17) {
          const double goq = gq * oos, gop = gp * oos;
          double v = 0.0;
          switch (t) {
            case 0: v = PA.x; break;
            case 1: v = PA.y; break;
            case 2: v = PA.z; break;
            case 3: v = QC.x; break;
            case 4: v = QC.y; break;
            case 5: v = QC.z; break;
            case 6: v = -goq * PQx; break;
            case 7: v = -goq * PQy; break;
            case 8: v = -goq * PQz; break;
            case 9: v = gop * PQx; break;
            case 10: v = gop * PQy; break;
            case 11: v = gop * PQz; break;
            case 12: v = xb.y; break;   // 1/(2 zeta)
            case 13: v = xk.y; break;   // 1/(2 eta)
            case 14: v = goq; break;    // rho/zeta
            case 15: v = gop; break;    // rho/eta
            default: v = 0.5 * oos;     // 1/(2 (zeta + eta))
          }
          lds[C.coef + so + t] = v;
        }
        if (t < C.m0len) {
          // erf_coulomb (boys.h:1913-1944): G_m = s^(m+1/2) F_m(s T), s = w^2/(w^2 + rho)
          if (A.omega > 0) {
            const double w2 = A.omega * A.omega, s = w2 / (w2 + gp * gq * oos);
            lds[C.base + so + t] = pf * pow(s, t + 0.5) * boys_one(T * s, t, A.boys);
          } else {
            lds[C.base + so + t] = pf * boys_one(T, t, A.boys);
          }
        }
      }
      if (ns == 0) break;
      const bool two = ns == 2;
      const int s2 = C.pstride;
      __syncthreads();
      // levels; a thread's next record is loaded (at a clamped, always valid
      // index) while it computes the current one, and its first record of
      // the next level before the barrier
      int nd = 0, nend = 0;  // the class's level table has L + 1 entries
      if (C.L > 0) {
        nd = idx[C.lev_first] + t;
        nend = idx[C.lev_first + 1];
      }
      int4 rc = prog[C.vrr_first + min(nd, max(nend - 1, 0))];
      for (int k = 1; k <= C.L; ++k) {
        const int nnext = k < C.L ? idx[C.lev_first + k + 1] : 0;
        const int cnt = nd < nend ? (nend - nd + kThreads - 1) / kThreads : 0;
        for (int i = 0; i < cnt; ++i) {
          const int ndn = nd + kThreads;
          const int4 rn = prog[C.vrr_first + min(ndn, nend - 1)];
          const int out = rc.x & 0x7FFF, mlen = (rc.x >> 15) & 15, slot = (rc.x >> 19) & 7;
          const int na = (rc.x >> 22) & 7, nc = (rc.x >> 25) & 7, kind = (rc.x >> 28) & 1;
          const double cA = lds[C.coef + slot], cW = lds[C.coef + 6 + slot];
          const double g = (double)na * lds[C.coef + 12 + kind], r = lds[C.coef + 14 + kind];
          const double h = (double)nc * lds[C.coef + 16];
          const double* X = lds + (rc.y & 0xFFFF);
          const double* Y = lds + (rc.y >> 16);
          const double* Z = lds + (rc.z & 0xFFFF);
          double* O = lds + out;
          // X[m] and Y[m] carried from the previous m (three LDS reads per
          // value, not five): O lies in level k's region, X, Y and Z in
          // levels k - 1 and k - 2 (or the zero vector), so no store of O
          // changes a value read later
          const int tg = (rc.z >> 16) - 1;
          if (two) {
            const double cA1 = lds[C.coef + s2 + slot], cW1 = lds[C.coef + s2 + 6 + slot];
            const double g1 = (double)na * lds[C.coef + s2 + 12 + kind], r1 = lds[C.coef + s2 + 14 + kind];
            const double h1 = (double)nc * lds[C.coef + s2 + 16];
            double x0 = X[0], y0 = Y[0], u0 = X[s2], v0 = Y[s2];
#pragma unroll 2
            for (int m = 0; m < mlen; ++m) {
              const double x1 = X[m + 1], y1 = Y[m + 1], u1 = X[s2 + m + 1], v1 = Y[s2 + m + 1];
              O[m] = cA * x0 + cW * x1 + g * (y0 - r * y1) + h * Z[m + 1];
              O[s2 + m] = cA1 * u0 + cW1 * u1 + g1 * (v0 - r1 * v1) + h1 * Z[s2 + m + 1];
              x0 = x1, y0 = y1, u0 = u1, v0 = v1;
            }
            if (tg >= 0) {
              lds[C.acc + tg] += O[0];
              lds[C.acc + tg] += O[s2];
            }
          } else {
            double x0 = X[0], y0 = Y[0];
#pragma unroll 2
            for (int m = 0; m < mlen; ++m) {
              const double x1 = X[m + 1], y1 = Y[m + 1];
              O[m] = cA * x0 + cW * x1 + g * (y0 - r * y1) + h * Z[m + 1];
              x0 = x1, y0 = y1;
            }
            if (tg >= 0) lds[C.acc + tg] += O[0];
          }
          rc = rn;
          nd = ndn;
        }
        // first record of level k + 1 for this thread
        nd = nend + t;
        nend = nnext;
        if (k < C.L) rc = prog[C.vrr_first + min(nd, max(nend - 1, 0))];
        __syncthreads();
      }
      if (C.targ0 >= 0 && t == 0) {
        lds[C.acc + C.targ0] += lds[C.base];
        if (two) lds[C.acc + C.targ0] += lds[C.base + s2];
      }
    }
    __syncthreads();
    if (gpart == 1) {
      // a piece: its contracted accumulators and survivor count to its row
      double* row = A.gpart + (long long)qi * A.gpart_stride;
      const int nacc = ne * C.accs;
      for (int i = t; i < nacc; i += kThreads) row[i] = lds[C.acc + i];
      if (t == 0) row[nacc] = (double)nsurv;
      continue;  // block-uniform
    }
    if (MODE == kModeEri && t == 0 && A.nz) A.nz[A.qidx[qi]] = nsurv > 0 ? 1 : 0;
    if (MODE == kModeFock && nsurv == 0) continue;  // nothing to digest (libint's nullptr)
    double smax = 0.0, ssq = 0.0, hmax = 0.0;
    if (MODE == kModeSchwarz)
      for (int i = t; i < C.ndiag; i += kThreads) hmax = fmax(hmax, fabs(lds[C.acc + idx[C.diag_first + i]]));
    // ---- 2. ket HRR and ket solid harmonics, one thread per bra row
    const int pk1 = (meta >> 5) & 1, pk2 = (meta >> 6) & 1;
    const int nc3 = ncart_d(lc), nc4 = ncart_d(ld), ncc = nc3 * nc4;
    const int nk1 = pk1 && lc > 0 ? 2 * lc + 1 : nc3, nk2 = pk2 && ld > 0 ? 2 * ld + 1 : nc4;
    for (int r0 = 0; r0 < ne; r0 += C.chk) {
      const int ie = r0 + t;
      if (t < C.chk && ie < ne) {
        double* ws = lds + C.w0 + t * C.wrow;
        double* tmp = ws + C.hk_nws;
        for (int j = 0; j < nf; ++j) ws[j] = lds[C.acc + ie * C.accs + j];
        for (int o = 0; o < C.hk_nops; ++o) {
          const int4 op = prog[C.hk_first + o];
          ws[op.x] = ws[op.y] + CD[op.w] * ws[op.z];
        }
        for (int k = 0; k < ncc; ++k) tmp[k] = ws[idx[C.hk_out + k]];
        sph_first(tmp, ws, lc, nc4, pk1);
        sph_second(ws, lds + C.y0 + ie * C.ys, nk1, ld, pk2);
      }
    }
    __syncthreads();
    // ---- 3. bra HRR, bra solid harmonics and the store / digest / norm
    const int pb1 = (meta >> 3) & 1, pb2 = (meta >> 4) & 1;
    const int nc1 = ncart_d(la), nc2 = ncart_d(lb), nab = nc1 * nc2;
    const int n1 = pb1 && la > 0 ? 2 * la + 1 : nc1, n2 = pb2 && lb > 0 ? 2 * lb + 1 : nc2;
    const int ncol = nk1 * nk2;
    L2gStore st;
    st.gauss = A.shg_gaussian;
    st.l1 = la, st.l2 = lb, st.l3 = lc, st.l4 = ld;
    st.p1 = pb1, st.p2 = pb2, st.p3 = pk1, st.p4 = pk2;
    st.n1 = n1, st.n2 = n2, st.nk1 = nk1, st.nk2 = nk2;
    st.uniform = A.uniform_cart;
    if (MODE == kModeEri) st = make_store(meta, A, qi, la, lb, lc, ld);
    // Fock: matrices of this item's request, accumulators J12 [n1 n2],
    // K13 [n1 nk1], K14 [n1 nk2], K24 [n2 nk2], K23 [n2 nk1]
    long long N = 0, moff = 0;
    int4 bf = int4{0, 0, 0, 0};
    double deg = 0.0;
    double *dJ12 = nullptr, *dK13 = nullptr, *dK14 = nullptr, *dK24 = nullptr, *dK23 = nullptr;
    const bool fx = MODE == kModeFock && A.fx_hoff != 0;  // deterministic (fixed-point) digest
    int ndg = 0;
    if (MODE == kModeFock) {
      const L2gMatDesc md = A.mats[A.q_req ? A.q_req[qi] : 0];
      N = md.n;
      moff = md.off;
      bf = A.q_bf[qi];
      deg = A.q_scale[qi];
      dJ12 = lds + C.f0;
      dK13 = dJ12 + n1 * n2;
      dK14 = dK13 + n1 * nk1;
      dK24 = dK14 + n1 * nk2;
      dK23 = dK24 + n2 * nk2;
      ndg = n1 * n2 + (n1 + n2) * (nk1 + nk2);
      // fixed-point accumulation: low words here, high words ndg further on
      for (int i = t; i < (fx ? 2 * ndg : ndg); i += kThreads) dJ12[i] = 0.0;
      __syncthreads();
    }
    const double* Dm = A.dens + moff;
    for (int c0 = 0; c0 < ncol; c0 += C.chb) {
      const int col = c0 + t;
      if (t < C.chb && col < ncol) {
        double* ws = lds + C.wb0 + t * C.wcol;
        double* tmp = ws + C.hb_nws;
        for (int j = 0; j < ne; ++j) ws[j] = lds[C.y0 + j * C.ys + col];
        for (int o = 0; o < C.hb_nops; ++o) {
          const int4 op = prog[C.hb_first + o];
          ws[op.x] = ws[op.y] + AB[op.w] * ws[op.z];
        }
        for (int k = 0; k < nab; ++k) tmp[k] = ws[idx[C.hb_out + k]];
        sph_first(tmp, ws, la, nc2, pb1);
        sph_second(ws, tmp, n1, lb, pb2);  // values v[i1 n2 + i2] in tmp
        const int ic = col / nk2, id = col - ic * nk2;
        if (MODE == kModeEri) {
          const size_t colofs = st.base + (size_t)st.pos3(ic) * st.s3 + (size_t)st.pos4(id) * st.s4;
          const double colscale = st.uniform ? st.scale3(ic) * st.scale4(id) : 1.0;
          for (int i1 = 0; i1 < n1; ++i1)
            for (int i2 = 0; i2 < n2; ++i2) st.put(A.out, colofs, i1, i2, colscale, tmp[i1 * n2 + i2]);
        } else if (MODE == kModeSchwarz) {
          for (int k = 0; k < n1 * n2; ++k) {
            smax = fmax(smax, fabs(tmp[k]));
            ssq += tmp[k] * tmp[k];
          }
        } else {
          const long long bk3 = bf.z + st.pos3(ic), bk4 = bf.w + st.pos4(id);
          const double d34 = Dm[bk3 * N + bk4];
          double jcol = 0.0;
          for (int i1 = 0; i1 < n1; ++i1) {
            const long long ba1 = bf.x + st.pos(la, pb1, i1);
            const double d13 = Dm[ba1 * N + bk3], d14 = Dm[ba1 * N + bk4];
            double k13 = 0.0, k14 = 0.0;
            for (int i2 = 0; i2 < n2; ++i2) {
              const long long ba2 = bf.y + st.pos(lb, pb2, i2);
              const double v = tmp[i1 * n2 + i2];
              jcol += Dm[ba1 * N + ba2] * v;
              lds_acc(&dJ12[i1 * n2 + i2], d34 * v, fx, ndg);
              k13 += Dm[ba2 * N + bk4] * v;
              k14 += Dm[ba2 * N + bk3] * v;
              lds_acc(&dK24[i2 * nk2 + id], d13 * v, fx, ndg);
              lds_acc(&dK23[i2 * nk1 + ic], d14 * v, fx, ndg);
            }
            lds_acc(&dK13[i1 * nk1 + ic], k13, fx, ndg);
            lds_acc(&dK14[i1 * nk2 + id], k14, fx, ndg);
          }
          l2g_acc(&A.jg[moff + bk3 * N + bk4], deg * jcol, A.fx_hoff);
        }
      }
    }
    if (MODE == kModeFock) {
      __syncthreads();
      double* jg = A.jg + moff;
      double* kg = A.kg + moff;
      const long long hoff = A.fx_hoff;
      auto val = [&](const double* p) { return lds_value(p, fx, ndg); };
      for (int i = t; i < n1 * n2; i += kThreads) {
        const int i1 = i / n2, i2 = i - i1 * n2;
        l2g_acc(&jg[(bf.x + st.pos(la, pb1, i1)) * N + bf.y + st.pos(lb, pb2, i2)], deg * val(&dJ12[i]), hoff);
      }
      for (int i = t; i < n1 * nk1; i += kThreads) {
        const int i1 = i / nk1, ic = i - i1 * nk1;
        l2g_acc(&kg[(bf.x + st.pos(la, pb1, i1)) * N + bf.z + st.pos3(ic)], deg * val(&dK13[i]), hoff);
      }
      for (int i = t; i < n1 * nk2; i += kThreads) {
        const int i1 = i / nk2, id = i - i1 * nk2;
        l2g_acc(&kg[(bf.x + st.pos(la, pb1, i1)) * N + bf.w + st.pos4(id)], deg * val(&dK14[i]), hoff);
      }
      for (int i = t; i < n2 * nk2; i += kThreads) {
        const int i2 = i / nk2, id = i - i2 * nk2;
        l2g_acc(&kg[(bf.y + st.pos(lb, pb2, i2)) * N + bf.w + st.pos4(id)], deg * val(&dK24[i]), hoff);
      }
      for (int i = t; i < n2 * nk1; i += kThreads) {
        const int i2 = i / nk1, ic = i - i2 * nk1;
        l2g_acc(&kg[(bf.y + st.pos(lb, pb2, i2)) * N + bf.z + st.pos3(ic)], deg * val(&dK23[i]), hoff);
      }
    }
    if (MODE == kModeSchwarz) {
      // block reductions of max|v|, sum v^2 and the pre-HRR diagonal max
      for (int o = 32; o > 0; o >>= 1) {
        smax = fmax(smax, __shfl_xor(smax, o));
        ssq += __shfl_xor(ssq, o);
        hmax = fmax(hmax, __shfl_xor(hmax, o));
      }
      const int w = t >> 6, nw = kThreads >> 6;
      if ((t & 63) == 0) {
        red[w] = smax;
        red[nw + w] = ssq;
      }
      __syncthreads();
      if (t == 0) {
        double m = 0.0, q = 0.0;
        for (int i = 0; i < nw; ++i) m = fmax(m, red[i]), q += red[nw + i];
        A.out[A.out_off[qi]] = A.schwarz_norm ? sqrt(sqrt(q)) : sqrt(m);
      }
      __syncthreads();
      if ((t & 63) == 0) red[w] = hmax;
      __syncthreads();
      if (t == 0 && A.out2) {
        double m = 0.0;
        for (int i = 0; i < nw; ++i) m = fmax(m, red[i]);
        A.out2[A.out_off[qi]] = sqrt(m);
      }
    }
    __syncthreads();  // LDS is reused by the next quartet
  }
}

// ---------------------------------------------------------------------------
// One-body integrals: overlap, kinetic energy and point-charge attraction
// over contracted Cartesian shells, then the solid-harmonic transform of
// pure shells -- what Engine(Operator::overlap / kinetic / nuclear) computes
// per shell pair (engine.impl.h:181-561) and the reference's driver loops
// over (tests/hartree-fock/hartree-fock.cc:480-535).  McMurchie-Davidson
// (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure Theory",
// 9.5, 9.7, 9.9): Hermite coefficients E^{ij}_t of the Gaussian product per
// axis (j to lb + 2 for the kinetic operator -1/2 nabla^2 on the ket),
// Hermite Coulomb integrals R_{tuv} per charge.  A 64-thread block per shell
// pair (blockIdx.x = bra, blockIdx.y = ket); a thread owns up to two of the
// at most 100 Cartesian pairs; the E and R tables sit in LDS, built by one
// thread per axis / charge.  Coefficients are used as the basis holds them.
// (t, u, v) with t + u + v <= 6 ordered by t + u + v: the entries of
// total <= k are the first (k+1)(k+2)(k+3)/6
struct L2gTuv {
  signed char v[84][3];
};
constexpr L2gTuv make_tuv() {
  L2gTuv r{};
  int e = 0;
  for (int s = 0; s <= 6; ++s)
    for (int t = s; t >= 0; --t)
      for (int u = s - t; u >= 0; --u) {
        r.v[e][0] = (signed char)t;
        r.v[e][1] = (signed char)u;
        r.v[e][2] = (signed char)(s - t - u);
        ++e;
      }
  return r;
}
__constant__ signed char c_tuv[84][3] = {};

__global__ void __launch_bounds__(64) l2g_onebody_kernel(l2g_impl::L2gOneBodyArgs A) {
  // blocks with bra >= ket; the ket > bra blocks are their transposes, as
  // the reference's driver fills them (hartree-fock.cc:510-530)
  const int sa = blockIdx.x, sb = blockIdx.y;
  if (sa >= A.nshell || sb >= A.nshell || sb > sa) return;
  const int la = A.l[sa], lb = A.l[sb];
  const int nca = ncart_d(la), ncb = ncart_d(lb), npair = nca * ncb;
  const int tid = threadIdx.x;
  __shared__ double E[3][4][6][10];   // [axis][i <= 3][j <= 5][t <= 9]
  __shared__ double Rw[7][7][7][7];   // [n][t][u][v], t + u + v + n <= la + lb <= 6
  __shared__ double cart[3][100];     // contracted Cartesian S, T, V
  int ex[2][3], ey[2][3];  // exponents of the owned bra / ket components
  bool own[2];
  for (int k = 0; k < 2; ++k) {
    const int r = tid + 64 * k;
    own[k] = r < npair;
    const int ca = own[k] ? r / ncb : 0, cb = own[k] ? r % ncb : 0;
    for (int side = 0; side < 2; ++side) {
      const int l = side ? lb : la, c = side ? cb : ca;
      int rr = 0;
      while ((rr + 1) * (rr + 2) / 2 <= c) ++rr;
      int* e = side ? ey[k] : ex[k];
      e[2] = c - rr * (rr + 1) / 2;
      e[1] = rr - e[2];
      e[0] = l - rr;
    }
  }
  double accS[2] = {0.0, 0.0}, accT[2] = {0.0, 0.0}, accV[2] = {0.0, 0.0};
  const double* Ao = A.origin + 3 * sa;
  const double* Bo = A.origin + 3 * sb;
  const int L = la + lb;
  for (int ia = A.poff[sa]; ia < A.poff[sa + 1]; ++ia)
    for (int ib = A.poff[sb]; ib < A.poff[sb + 1]; ++ib) {
      const double a = A.alpha[ia], b = A.alpha[ib], p = a + b, cc = A.coeff[ia] * A.coeff[ib];
      double P[3];
      for (int k = 0; k < 3; ++k) P[k] = (a * Ao[k] + b * Bo[k]) / p;
      if (tid < 3) {
        const int k = tid;
        const double XAB = Ao[k] - Bo[k], XPA = -b / p * XAB, XPB = a / p * XAB, o2p = 0.5 / p;
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 6; ++j)
            for (int t = 0; t < 10; ++t) E[k][i][j][t] = 0.0;
        E[k][0][0][0] = exp(-a * b / p * XAB * XAB);
        for (int i = 0; i <= la; ++i)
          for (int j = 0; j <= lb + 2; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
              double v;
              if (i > 0)
                v = XPA * E[k][i - 1][j][t] + (t > 0 ? o2p * E[k][i - 1][j][t - 1] : 0.0) +
                    (t + 1) * E[k][i - 1][j][t + 1];
              else
                v = XPB * E[k][i][j - 1][t] + (t > 0 ? o2p * E[k][i][j - 1][t - 1] : 0.0) +
                    (t + 1) * E[k][i][j - 1][t + 1];
              E[k][i][j][t] = v;
            }
          }
      }
      __syncthreads();
      const double s1 = sqrt(3.14159265358979323846 / p);
      for (int k = 0; k < 2; ++k) {
        if (!own[k]) continue;
        double ov[3][7];  // ov[axis][j'] for j' = j - 2 .. j + 2 at index j' - j + 2
        for (int x = 0; x < 3; ++x)
          for (int d = -2; d <= 2; ++d) {
            const int j = ey[k][x] + d;
            ov[x][d + 2] = j >= 0 ? E[x][ex[k][x]][j][0] * s1 : 0.0;
          }
        accS[k] += cc * ov[0][2] * ov[1][2] * ov[2][2];
        double t = 0.0;
        for (int x = 0; x < 3; ++x) {
          const int j = ey[k][x];
          double tk = b * (2 * j + 1) * ov[x][2] - 2.0 * b * b * ov[x][4];
          if (j >= 2) tk -= 0.5 * j * (j - 1) * ov[x][0];
          t += tk * ov[(x + 1) % 3][2] * ov[(x + 2) % 3][2];
        }
        accT[k] += cc * t;
      }
      for (int c = 0; c < A.ncharge; ++c) {
        // R^n_{000} = (-2p)^n F_n(T), one n per thread; then level by level
        // n = L-1 .. 0, the entries 0 < t+u+v <= L-n of a level spread over
        // the threads (each level reads only the one above)
        const double PC[3] = {P[0] - A.qxyz[3 * c], P[1] - A.qxyz[3 * c + 1], P[2] - A.qxyz[3 * c + 2]};
        if (tid <= L) {
          const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
          double f = 1.0;
          for (int n = 0; n < tid; ++n) f *= -2.0 * p;
          Rw[tid][0][0][0] = f * boys_one(T, tid, A.boys);
        }
        __syncthreads();
        for (int n = L - 1; n >= 0; --n) {
          const int top = L - n, cnt = (top + 1) * (top + 2) * (top + 3) / 6;  // t+u+v <= top
          for (int e = tid + 1; e < cnt; e += 64) {
            const int t = c_tuv[e][0], u = c_tuv[e][1], v = c_tuv[e][2];
            double val;
            if (t > 0)
              val = PC[0] * Rw[n + 1][t - 1][u][v] + (t > 1 ? (t - 1) * Rw[n + 1][t - 2][u][v] : 0.0);
            else if (u > 0)
              val = PC[1] * Rw[n + 1][t][u - 1][v] + (u > 1 ? (u - 1) * Rw[n + 1][t][u - 2][v] : 0.0);
            else
              val = PC[2] * Rw[n + 1][t][u][v - 1] + (v > 1 ? (v - 1) * Rw[n + 1][t][u][v - 2] : 0.0);
            Rw[n][t][u][v] = val;
          }
          __syncthreads();
        }
        const double zf = A.q[c] * 2.0 * 3.14159265358979323846 / p;
        for (int k = 0; k < 2; ++k) {
          if (!own[k]) continue;
          double acc = 0.0;
          for (int t = 0; t <= ex[k][0] + ey[k][0]; ++t)
            for (int u = 0; u <= ex[k][1] + ey[k][1]; ++u)
              for (int v = 0; v <= ex[k][2] + ey[k][2]; ++v)
                acc += E[0][ex[k][0]][ey[k][0]][t] * E[1][ex[k][1]][ey[k][1]][u] * E[2][ex[k][2]][ey[k][2]][v] *
                       Rw[0][t][u][v];
          accV[k] -= cc * zf * acc;
        }
        __syncthreads();
      }
      __syncthreads();
    }
  for (int k = 0; k < 2; ++k)
    if (own[k]) {
      cart[0][tid + 64 * k] = accS[k];
      cart[1][tid + 64 * k] = accT[k];
      cart[2][tid + 64 * k] = accV[k];
    }
  __syncthreads();
  const int pa = (A.pure[sa] && la > 0) ? 1 : 0, pb = (A.pure[sb] && lb > 0) ? 1 : 0;
  const int noa = pa ? 2 * la + 1 : nca, nob = pb ? 2 * lb + 1 : ncb;
  double* out[3] = {A.S, A.T, A.V};
  for (int r = tid; r < noa * nob; r += 64) {
    const int fi = r / nob, fj = r % nob;
    // output position -> row of c_sh (m + l): Standard m = f - l, Gaussian 0, +1, -1, +2, ...
    const int ra = pa ? (A.gauss ? la + (fi == 0 ? 0 : ((fi & 1) ? (fi + 1) / 2 : -fi / 2)) : fi) : 0;
    const int rb = pb ? (A.gauss ? lb + (fj == 0 ? 0 : ((fj & 1) ? (fj + 1) / 2 : -fj / 2)) : fj) : 0;
    const long long o = (A.bf0[sa] + fi) * A.nbf + A.bf0[sb] + fj;
    const long long ot = (A.bf0[sb] + fj) * A.nbf + A.bf0[sa] + fi;
    for (int m = 0; m < 3; ++m) {
      if (!out[m]) continue;
      double s = 0.0;
      for (int ca = pa ? 0 : fi; ca < (pa ? nca : fi + 1); ++ca) {
        const double wa = pa ? c_sh[la][ra][ca] : 1.0;
        if (wa == 0.0) continue;
        for (int cb = pb ? 0 : fj; cb < (pb ? ncb : fj + 1); ++cb) {
          const double wb = pb ? c_sh[lb][rb][cb] : 1.0;
          if (wb != 0.0) s += wa * wb * cart[m][ca * ncb + cb];
        }
      }
      out[m][o] = s;
      if (sa != sb) out[m][ot] = s;  // a diagonal block writes each of its elements itself
    }
  }
}

}  // namespace

namespace l2g_impl {

void launch_onebody(const L2gOneBodyArgs& A, hipStream_t s) {
  if (A.nshell <= 0) return;
  hipLaunchKernelGGL(l2g_onebody_kernel, dim3(A.nshell, A.nshell), dim3(64), 0, s, A);
  LAUNCHCHK("l2g_onebody_kernel");
}

template <int MODE>
void set_lds_limit(int maxlds) {
  HIPCHK(hipFuncSetAttribute((const void*)l2g_generic_kernel<MODE, 256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             maxlds));
  HIPCHK(hipFuncSetAttribute((const void*)l2g_generic_kernel<MODE, 512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             maxlds));
  HIPCHK(hipFuncSetAttribute((const void*)l2g_generic_kernel<MODE, 1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             maxlds));
}

template <int MODE>
void launch_mode(const L2gEriArgs& A, const L2gGenericClass& C, const int4* prog, const int* idx, int grid,
                 hipStream_t s) {
  const size_t shm = (size_t)C.lds * sizeof(double);
  const size_t per_cu = (160 * 1024) / (shm + 512);
  if (per_cu <= 1)
    hipLaunchKernelGGL((l2g_generic_kernel<MODE, 1024>), dim3(grid), dim3(1024), shm, s, A, C, prog, idx);
  else if (per_cu <= 3)
    hipLaunchKernelGGL((l2g_generic_kernel<MODE, 512>), dim3(grid), dim3(512), shm, s, A, C, prog, idx);
  else
    hipLaunchKernelGGL((l2g_generic_kernel<MODE, 256>), dim3(grid), dim3(256), shm, s, A, C, prog, idx);
}

void generic_init(const double* sh) {
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_sh), sh, sizeof(double) * 4 * 7 * 10));
  constexpr L2gTuv tuv = make_tuv();  // one-body kernel's Hermite index table
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_tuv), tuv.v, sizeof(tuv.v)));
  // 160 KB per CU on gfx950, less 512 bytes for the static reduction array
  // (codegen/gen_generic.py plans at most that many doubles)
  const int maxlds = 160 * 1024 - 512;
  set_lds_limit<kModeEri>(maxlds);
  set_lds_limit<kModeFock>(maxlds);
  set_lds_limit<kModeSchwarz>(maxlds);
}

void launch_generic(int mode, const L2gEriArgs& A, const L2gGenericClass& C, const int4* prog, const int* idx,
                    int grid, hipStream_t s) {
  if (mode == kModeEri)
    launch_mode<kModeEri>(A, C, prog, idx, grid, s);
  else if (mode == kModeFock)
    launch_mode<kModeFock>(A, C, prog, idx, grid, s);
  else
    launch_mode<kModeSchwarz>(A, C, prog, idx, grid, s);
  LAUNCHCHK("generic kernel");
}

}  // namespace l2g_impl
