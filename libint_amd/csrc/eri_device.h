// Device-side definitions shared by the generated ERI kernels and the host
// launcher: kernel argument block, Boys function, permuted store.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define L2G_BOYS_NINT 820   // 7 intervals per unit T on [0, 117 + 1/7)
#define L2G_BOYS_MMAX 40    // tabulated m range, as FmEval_Chebyshev7 (boys.h:298-306)
#define L2G_BOYS_TCRIT 117.0

#define L2G_SCREEN_ORIGINAL 0x0001
#define L2G_SCREEN_CONSERVATIVE 0x0010

// Work list of one canonical class plus everything a kernel reads.
// Primitive-pair records are stored in the canonical orientation of the pair
// (first shell = the one with the larger angular momentum):
//   pp_P   = (Px, Py, Pz, gamma)                 gamma = alpha1 + alpha2
//   pp_PA  = (PAx, PAy, PAz, Kc)                 Kc = K * c1 * c2,
//            K = sqrt(2) pi^(5/4) exp(-rho |AB|^2) / gamma (shell.h:1244)
//   pp_aux = (ln_scr, 1 / (2 gamma))
//   pp_nonsph = conservative-screening factor (shell.h:1208-1227)
// Where one request's nbf x nbf matrices sit in the Fock arenas: D at
// dens + off, the jg / kg accumulators at jg + off / kg + off.
struct L2gMatDesc {
  long long off;
  long long n;
};

// Deterministic J/K accumulation.  A contribution v enters as the 128-bit
// fixed-point number floor(v 2^80), held as hi = floor(v 2^16) (two's
// complement) and lo = frac(v 2^16) 2^64.  Integer additions commute, so the
// accumulated pair does not depend on the order in which the contributions
// arrive: a build gives the same bits on every run and in every batch.
// Resolution 2^-80 per contribution; the sum must stay below 2^37 in
// magnitude for the conversion back to be exact in its high word.
__device__ __forceinline__ void l2g_fx_split(double v, unsigned long long& hi, unsigned long long& lo) {
  const double y = v * 65536.0;  // exact: a power of two
  const double h = floor(y);
  hi = (unsigned long long)(long long)h;
  lo = (unsigned long long)((y - h) * 18446744073709551616.0);  // y - h is exact, in [0, 1)
}

// (hi, lo) += v, with the carry out of the low word into the high word
__device__ __forceinline__ void l2g_fx_add(unsigned long long* hi, unsigned long long* lo, double v) {
  unsigned long long h, l;
  l2g_fx_split(v, h, l);
  const unsigned long long old = atomicAdd(lo, l);
  h += (old + l < old) ? 1ull : 0ull;
  if (h) atomicAdd(hi, h);
}

__device__ __forceinline__ double l2g_fx_value(unsigned long long hi, unsigned long long lo) {
  return (double)(long long)hi * (1.0 / 65536.0) + (double)lo * 0x1p-80;
}

// J/K accumulator update: an FP64 atomic add (hoff == 0), or the fixed-point
// pair whose low word is the 64 bits at p and whose high word is hoff words
// further on (a second plane of the accumulator arena)
__device__ __forceinline__ void l2g_acc(double* p, double v, long long hoff) {
  if (hoff == 0) {
    unsafeAtomicAdd(p, v);
    return;
  }
  unsigned long long* lo = reinterpret_cast<unsigned long long*>(p);
  l2g_fx_add(lo + hoff, lo, v);
}

// sum over the 64 lanes of the wave (butterfly), returned in every lane
template <class T>
__device__ __forceinline__ T l2g_wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct L2gEriArgs {
  const int* __restrict__ bra_pair;
  const int* __restrict__ ket_pair;
  const long long* __restrict__ out_off;
  const int* __restrict__ meta;  // bit0 swap_tbra, bit1 swap_tket, bit2 swap_braket,
                                 // bits3..6 pure flags of canonical b1,b2,k1,k2
  int n;
  const int* __restrict__ pair_first;
  const int* __restrict__ pair_np;
  const double* __restrict__ pair_AB;
  const double4* __restrict__ pp_P;
  const double4* __restrict__ pp_PA;
  const double2* __restrict__ pp_aux;
  const double* __restrict__ pp_nonsph;
  const double* __restrict__ boys;  // [m][iv][8]
  double precision, ln_precision, scale;
  int method;
  int shg_gaussian;
  int uniform_cart;
  int schwarz_norm;  // Schwarz kernels: 0 = max-abs norm, 1 = Frobenius norm
  int coop;          // Schwarz kernels: one wave per pair, its lanes splitting the primitive loop
  double omega;      // > 0: erf-attenuated Coulomb erf(omega r)/r (0: Coulomb)
  int store_sub;     // ERI kernels: subtract from the output instead of storing
  double* __restrict__ out;
  double* __restrict__ ybuf;
  long long ystride;
  signed char* __restrict__ nz;          // optional: 0 if every primitive was screened out
  const long long* __restrict__ qidx;    // caller index of each work item (for nz)
  // Fock digestion (l2g_fock_* kernels)
  const double* __restrict__ q_lnprec;   // per work item ln(precision)
  const double* __restrict__ q_scale;    // per work item degeneracy factor
  const int4* __restrict__ q_bf;         // first basis function of canonical b1, b2, k1, k2
  const double* __restrict__ dens;       // density arena (row-major nbf x nbf blocks)
  double* __restrict__ jg;               // unsymmetrised Coulomb accumulator arena
  double* __restrict__ kg;               // unsymmetrised exchange accumulator arena
  const L2gMatDesc* __restrict__ mats;   // matrix block of each request
  const int* __restrict__ q_req;         // request of each work item (null: request 0)
  const int2* __restrict__ q_prange;     // primitive quartets [x, y) of each work item (null: all)
  long long fx_hoff;                     // 0: FP64 atomics into jg / kg; else fixed point (l2g_acc)
  // Schwarz kernels: optional sqrt(max |[e0|e0]|) of the pre-HRR diagonal
  double* __restrict__ out2;
  // generic kernel, ERI mode, quartets split over blocks by primitive
  // range: 1 = a piece pass (work item = piece, q_prange its primitive
  // quartets; writes its contracted [e0|f0] accumulators and survivor
  // count to row gpart[item * gpart_stride]); 2 = the finish pass (work
  // item = quartet; sums rows gpart_first[item] .. gpart_first[item + 1] - 1
  // in order, then HRR and store); 0 = neither
  int gpart_mode;
  long long gpart_stride;
  double* __restrict__ gpart;
  const int* __restrict__ gpart_first;
};

typedef void (*L2gLaunchFn)(const L2gEriArgs&, int grid, int block, hipStream_t s);

// One class of the generic kernel (l2g_generic.hip): offsets into the
// program arrays written by codegen/gen_generic.py and the LDS layout of a
// block, all in doubles of LDS or in entries of the program / index arrays.
struct L2gGenericClass {
  int la, lb, lc, ld, L;
  int ne, nf;        // bra / ket components of the [e0|f0] targets
  int m0len, targ0;  // m-vector length of (00|00); its target index (-1 unless (ss|ss))
  int base, zero, coef, acc;
  int y0, w0, wrow, chk;  // ket phase: rows, row workspaces, rows per pass
  int f0, wb0, wcol, chb; // bra phase: digest accumulators, column workspaces, columns per pass
  int lds;                // doubles of dynamic LDS per block
  int vrr_first, lev_first;
  int hk_first, hk_nops, hk_out, hk_nws;
  int hb_first, hb_nops, hb_out, hb_nws;
  int diag_first, ndiag;
  int nvals;              // recursion values per primitive quartet
  int accs, ys;           // row strides of the accumulators and of Y (odd)
  int dual, pstride;      // two primitive quartets per recursion pass; offset of the second's block
};

struct L2gClassInfo {
  int la, lb, lc, ld;
  int ngroups;
  int ybuf_doubles;  // per lane; 0 = rows kept in registers
  int nodes;         // VRR graph nodes per primitive quartet (all groups)
  int fp32_pays;     // FP32 variants: measured faster than the class's FP64 kernel
  int few_generic;   // FP64 ERI kernels: few deep quartets measured faster on the generic kernel
  L2gLaunchFn launch;
};

// F_m(T) * pfac for m = 0..L.  Below T_crit: degree-7 Chebyshev interpolant
// of F_L on the interval holding T, then downward recursion
// F_{m-1} = (2T F_m + e^{-T}) / (2m - 1).  Above T_crit: the asymptotic
// F_0 = sqrt(pi/(4T)) and upward recursion without the e^{-T} term, as
// FmEval_Chebyshev7::eval does (boys.h:348-358).
template <int L>
__device__ __forceinline__ void boys_fm(double T, const double* __restrict__ tabL, double pfac, double* F) {
  if (T > L2G_BOYS_TCRIT) {
    const double oox = 1.0 / T;
    double f = 0.88622692545275801365 * sqrt(oox);
    F[0] = f * pfac;
#pragma unroll
    for (int m = 1; m <= L; ++m) {
      f = f * ((double)m - 0.5) * oox;
      F[m] = f * pfac;
    }
  } else {
    const double x7 = T * 7.0;
    const int iv = (int)x7;
    const double xd = x7 - (double)iv - 0.5;
    const double4* c4 = reinterpret_cast<const double4*>(tabL + (size_t)iv * 8);
    const double4 lo = c4[0], hi = c4[1];
    double f = hi.w;
    f = fma(f, xd, hi.z);
    f = fma(f, xd, hi.y);
    f = fma(f, xd, hi.x);
    f = fma(f, xd, lo.w);
    f = fma(f, xd, lo.z);
    f = fma(f, xd, lo.y);
    f = fma(f, xd, lo.x);
    double fm[L + 1];
    fm[L] = f;
    if (L > 0) {
      const double e = exp(-T);
      const double t2 = T + T;
#pragma unroll
      for (int m = L; m > 0; --m) fm[m - 1] = (t2 * fm[m] + e) * (1.0 / (double)(2 * m - 1));
    }
#pragma unroll
    for (int m = 0; m <= L; ++m) F[m] = fm[m] * pfac;
  }
}

// Boys function split in two steps so that the table row of the next
// primitive quartet can be fetched while the current one is being processed:
// boys_fetch locates the interval and loads its 8 coefficients (for T above
// T_crit the load is of a valid row but unused), boys_eval evaluates.
struct L2gBoysArg {
  double T, xd;
  double4 lo, hi;
};

__device__ __forceinline__ void boys_fetch(double T, const double* __restrict__ tabL, L2gBoysArg& b) {
  const double x7 = T * 7.0;
  const double ivd = fmin(floor(x7), (double)(L2G_BOYS_NINT - 1));
  const int iv = (int)ivd;
  b.T = T;
  b.xd = x7 - ivd - 0.5;
  const double4* c4 = reinterpret_cast<const double4*>(tabL + (size_t)iv * 8);
  b.lo = c4[0];
  b.hi = c4[1];
}

template <int L>
__device__ __forceinline__ void boys_eval(const L2gBoysArg& b, double pfac, double* F) {
  if (b.T > L2G_BOYS_TCRIT) {
    const double oox = 1.0 / b.T;
    double f = 0.88622692545275801365 * sqrt(oox);
    F[0] = f * pfac;
#pragma unroll
    for (int m = 1; m <= L; ++m) {
      f = f * ((double)m - 0.5) * oox;
      F[m] = f * pfac;
    }
  } else {
    const double xd = b.xd;
    double f = b.hi.w;
    f = fma(f, xd, b.hi.z);
    f = fma(f, xd, b.hi.y);
    f = fma(f, xd, b.hi.x);
    f = fma(f, xd, b.lo.w);
    f = fma(f, xd, b.lo.z);
    f = fma(f, xd, b.lo.y);
    f = fma(f, xd, b.lo.x);
    double fm[L + 1];
    fm[L] = f;
    if (L > 0) {
      const double e = exp(-b.T);
      const double t2 = b.T + b.T;
#pragma unroll
      for (int m = L; m > 0; --m) fm[m - 1] = (t2 * fm[m] + e) * (1.0 / (double)(2 * m - 1));
    }
#pragma unroll
    for (int m = 0; m <= L; ++m) F[m] = fm[m] * pfac;
  }
}

__device__ __forceinline__ int l2g_ncart(int l) { return (l + 1) * (l + 2) / 2; }

// sqrt((2l-1)!! / ((2i-1)!! (2j-1)!! (2k-1)!!)) for Cartesian index c of
// level l (cartesian.h:96-109); used only with uniform normalization.
__device__ __forceinline__ double l2g_uniform_cart_scale(int l, int c) {
  // components in standard order; compute exponents from the index
  int x = l, rem = c;
  while (rem > l - x) {
    rem -= (l - x + 1);
    --x;
  }
  const int y = l - x - rem, z = l - x - y;
  double num = 1, den = 1;
  for (int i = 2 * l - 1; i > 1; i -= 2) num *= i;
  for (int i = 2 * x - 1; i > 1; i -= 2) den *= i;
  for (int i = 2 * y - 1; i > 1; i -= 2) den *= i;
  for (int i = 2 * z - 1; i > 1; i -= 2) den *= i;
  return sqrt(num / den);
}

// Caller-order store of one quartet.  Canonical index k (b1,b2,k1,k2) lands
// at caller position pos[k] (engine.impl.h:1181-1208 swaps undone as in
// :1987-2067); within a pure shell, standard index s (m = s - l) is moved to
// the Gaussian position when that ordering is selected (shgshell_ordering.h).
struct L2gStore {
  size_t base;
  int n1, n2, nk1, nk2;
  int s1, s2, s3, s4;
  int l1, l2, l3, l4;
  int p1, p2, p3, p4;
  int gauss, uniform;
  int sub;  // subtract from the output (erfc = Coulomb - erf, second pass)

  __device__ __forceinline__ int pos(int l, int pure, int s) const {
    if (!pure || !gauss) return s;
    const int m = s - l;
    return m == 0 ? 0 : (m > 0 ? 2 * m - 1 : -2 * m);
  }
  __device__ __forceinline__ int pos3(int i) const { return pos(l3, p3, i); }
  __device__ __forceinline__ int pos4(int i) const { return pos(l4, p4, i); }
  __device__ __forceinline__ double scale3(int i) const { return p3 ? 1.0 : l2g_uniform_cart_scale(l3, i); }
  __device__ __forceinline__ double scale4(int i) const { return p4 ? 1.0 : l2g_uniform_cart_scale(l4, i); }
  __device__ __forceinline__ void put(double* __restrict__ out, size_t colofs, int i1, int i2, double colscale,
                                      double v) const {
    double val = v;
    if (uniform) {
      val *= colscale * (p1 ? 1.0 : l2g_uniform_cart_scale(l1, i1)) * (p2 ? 1.0 : l2g_uniform_cart_scale(l2, i2));
    }
    double* o = out + colofs + (size_t)pos(l1, p1, i1) * s1 + (size_t)pos(l2, p2, i2) * s2;
    *o = sub ? *o - val : val;
  }
  // Staged store: the value goes to its final position within the quartet's
  // block, but in the lane's scratch column (yb[k * ystride], k counted from
  // `stage`); flush() then writes the block to `out` in address order.  A
  // lane produces its block one ket column at a time, so direct stores leave
  // every cache line of the block half-written until the last column, and
  // the lines are evicted and written back more than once.
  __device__ __forceinline__ void put_staged(double* yb, size_t ystride, int stage, size_t colofs, int i1, int i2,
                                             double colscale, double v) const {
    double val = v;
    if (uniform) {
      val *= colscale * (p1 ? 1.0 : l2g_uniform_cart_scale(l1, i1)) * (p2 ? 1.0 : l2g_uniform_cart_scale(l2, i2));
    }
    const size_t k = colofs - base + (size_t)pos(l1, p1, i1) * s1 + (size_t)pos(l2, p2, i2) * s2;
    yb[((size_t)stage + k) * ystride] = val;
  }
  __device__ __forceinline__ void flush(double* __restrict__ out, const double* yb, size_t ystride, int stage) const {
    const int n = n1 * n2 * nk1 * nk2;
    const double* src = yb + (size_t)stage * ystride;
    double* dst = out + base;
    if (sub)
      for (int k = 0; k < n; ++k) dst[k] -= src[(size_t)k * ystride];
    else
      for (int k = 0; k < n; ++k) dst[k] = src[(size_t)k * ystride];
  }
};

__device__ __forceinline__ L2gStore make_store(int meta, const L2gEriArgs& A, int qi, int l1, int l2, int l3,
                                               int l4) {
  L2gStore st;
  st.base = (size_t)A.out_off[qi];
  st.l1 = l1, st.l2 = l2, st.l3 = l3, st.l4 = l4;
  st.p1 = (meta >> 3) & 1;
  st.p2 = (meta >> 4) & 1;
  st.p3 = (meta >> 5) & 1;
  st.p4 = (meta >> 6) & 1;
  st.sub = A.store_sub;
  const int n1 = st.p1 ? 2 * l1 + 1 : l2g_ncart(l1);
  const int n2 = st.p2 ? 2 * l2 + 1 : l2g_ncart(l2);
  const int n3 = st.p3 ? 2 * l3 + 1 : l2g_ncart(l3);
  const int n4 = st.p4 ? 2 * l4 + 1 : l2g_ncart(l4);
  st.n1 = n1;
  st.n2 = n2;
  st.nk1 = n3;
  st.nk2 = n4;
  const int swap_tbra = meta & 1, swap_tket = (meta >> 1) & 1, swap_braket = (meta >> 2) & 1;
  int pos[4];
  if (!swap_braket) {
    pos[0] = swap_tbra ? 1 : 0;
    pos[1] = swap_tbra ? 0 : 1;
    pos[2] = swap_tket ? 3 : 2;
    pos[3] = swap_tket ? 2 : 3;
  } else {
    pos[0] = swap_tket ? 3 : 2;
    pos[1] = swap_tket ? 2 : 3;
    pos[2] = swap_tbra ? 1 : 0;
    pos[3] = swap_tbra ? 0 : 1;
  }
  int csz[4];
  csz[pos[0]] = n1;
  csz[pos[1]] = n2;
  csz[pos[2]] = n3;
  csz[pos[3]] = n4;
  const int cs[4] = {csz[1] * csz[2] * csz[3], csz[2] * csz[3], csz[3], 1};
  st.s1 = cs[pos[0]];
  st.s2 = cs[pos[1]];
  st.s3 = cs[pos[2]];
  st.s4 = cs[pos[3]];
  st.gauss = A.shg_gaussian;
  st.uniform = A.uniform_cart;
  return st;
}
