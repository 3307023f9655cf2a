// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of libint's two-electron Coulomb path, used as the checker
// for the HIP implementation and as the timed CPU baseline.  Nothing in the
// shipped library links or calls this file; only tests/, the smoke check and
// the CPU-baseline leg of the benchmark load it.
//
// What it restates (reference files, relative to the libint source tree):
//   * Shell::renorm                         include/libint2/shell.h:958-999
//   * Shell::update_max_ln_coeff            include/libint2/shell.h:1001-1011
//   * ShellPair::init (Original/Conservative) include/libint2/shell.h:1137-1256
//   * FmEval_Chebyshev7::eval               include/libint2/boys.h:345-454
//     (table contents are not in the reference tree — .MISSING_LARGE_BLOBS:1;
//      this file builds its own degree-7 Chebyshev table with the documented
//      parameters of include/libint2/util/boys_cheb.nb: 7 intervals per unit
//      T, T_crit = 117, m <= 40.  Bit-level table parity is therefore
//      UNPINNED; accuracy is pinned against mpmath goldens.)
//   * FmEval_Reference / FmEval_Reference2  include/libint2/boys.h:143-247
//   * Engine::compute2<coulomb,xx_xx,0>     include/libint2/engine.impl.h:1151-2113
//       canonical reorder                   :1181-1221
//       primitive-quartet loop + screening  :1306-1766
//       (ss|ss) shortcut                    :1786-1804
//       VRR (Obara-Saika on a then c)       src/bin/libint/vrr_11_twoprep_11.h:152-224
//       contraction                         src/bin/libint/dg.cc:1177-1189
//       HRR (Head-Gordon-Pople, AB then CD) src/bin/libint/hrr.h:290-330
//       cart->sph                           include/libint2/solidharmonics.h:114-240
//       un-permute                          include/libint2/engine.impl.h:1987-2067
//       uniform cartesian normalization     include/libint2/cartesian.h:117-133
//   * SolidHarmonicsCoefficients::coeff     include/libint2/solidharmonics.h:114-180
//   * brute-force check integral: McMurchie-Davidson in long double.  This is
//     deliberately a different algorithm from src/bin/test_eri/eri.h:121-385
//     (which expands in Cartesian binomials) so that it shares no steps with
//     the Obara-Saika restatement it checks.
//   * screened Fock build G = 2J - K        tests/hartree-fock/hartree-fock++.cc:1574-1772
//     (semantics pinned by the unscreened hartree-fock.cc:536-626)
//
// Build: see oracle/Makefile (g++ -O3 -march=native -fopenmp).

#include <algorithm>
#include <array>
#include <atomic>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <quadmath.h>
#endif

namespace orc {

// ----------------------------------------------------------------------------
// small math tables
// ----------------------------------------------------------------------------
static double fac(int n) {
  double r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
static long double facl(int n) {
  long double r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
// (k-1)!!, with (-1)!! = 0!! = 1
static double df_km1(int k) {
  double r = 1;
  for (int i = k - 1; i > 1; i -= 2) r *= i;
  return r;
}
static double binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  return fac(n) / (fac(k) * fac(n - k));
}
static int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static int nsize(int l, int pure) { return pure ? 2 * l + 1 : ncart(l); }

// Cartesian components of level l in libint's standard order
// (include/libint2/cgshell_ordering.h:125-130: x descending, then y).
static std::vector<std::array<int, 3>> cart_comps(int l) {
  std::vector<std::array<int, 3>> r;
  for (int i = l; i >= 0; --i)
    for (int j = l - i; j >= 0; --j) r.push_back({i, j, l - i - j});
  return r;
}
static int cart_index(int x, int y, int z) {
  const int l = x + y + z;
  return ((l - x + 1) * (l - x)) / 2 + l - x - y;
}

// ----------------------------------------------------------------------------
// Boys function
// ----------------------------------------------------------------------------
// MacLaurin series in long double; restates FmEval_Reference::eval
// (boys.h:153-178): F_m(T) = e^{-T} sum_k T^k / ((2m+1)(2m+3)...(2m+2k+1)).
static long double boys_series(long double T, int m) {
  long double denom = m + 0.5L;
  long double term = expl(-T) / (2 * denom);
  long double sum = term, old;
  const long double eps = std::numeric_limits<long double>::epsilon() / 10;
  do {
    denom += 1;
    old = term;
    term = old * T / denom;
    sum += term;
  } while (term > sum * eps || old < term);
  return sum;
}
// FmEval_Reference2 (boys.h:213-247): series below 117, erf + upward above.
static void boys_reference(long double T, int mmax, long double* F) {
  if (T < 117) {
    for (int m = 0; m <= mmax; ++m) F[m] = boys_series(T, m);
  } else {
    const long double t2 = 2 * T, et = expl(-T), st = sqrtl(T);
    F[0] = 0.886226925452758013649083741671L * erfl(st) / st;
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - et) / t2;
  }
}

// Chebyshev-7 table: 7 intervals per unit of T on [0, 117], m in [0, 40],
// 8 monomial coefficients in xd = 7T - iv - 1/2 per (interval, m) —
// the layout FmEval_Chebyshev7 indexes at boys.h:422-441.
struct Cheb7 {
  static constexpr int ORDER = 7, NCOEF = 8, MMAX = 40;
  static constexpr double TCRIT = 117.0, DELTA = 1.0 / 7.0;
  // 117 * 7 intervals plus one, so that T = 117 exactly (which is not above
  // T_crit) still finds its interval
  static constexpr int NINT = 820;
  std::vector<double> c;            // [NINT][MMAX+1][8]
  Cheb7() {
    c.resize((size_t)NINT * (MMAX + 1) * NCOEF);
    const long double PI = 3.141592653589793238462643383279502884L;
    // Chebyshev nodes x_k on [-1,1]; xd = x/2
    long double xk[NCOEF];
    for (int k = 0; k < NCOEF; ++k) xk[k] = cosl(PI * (k + 0.5L) / NCOEF);
    // monomial expansion of Chebyshev polynomials T_j(x) = sum_i tc[j][i] x^i
    long double tc[NCOEF][NCOEF] = {};
    tc[0][0] = 1;
    tc[1][1] = 1;
    for (int j = 2; j < NCOEF; ++j)
      for (int i = 0; i < NCOEF; ++i)
        tc[j][i] = (i > 0 ? 2 * tc[j - 1][i - 1] : 0) - tc[j - 2][i];
    std::vector<long double> F((MMAX + 1) * NCOEF);
    for (int iv = 0; iv < NINT; ++iv) {
      for (int k = 0; k < NCOEF; ++k) {
        const long double T = (iv + 0.5L + xk[k] / 2) / 7.0L;
        // series for the top m, then stable downward recursion
        long double fm[MMAX + 1];
        fm[MMAX] = boys_series(T, MMAX);
        const long double eT = expl(-T);
        for (int m = MMAX; m > 0; --m) fm[m - 1] = (2 * T * fm[m] + eT) / (2 * m - 1);
        for (int m = 0; m <= MMAX; ++m) F[m * NCOEF + k] = fm[m];
      }
      for (int m = 0; m <= MMAX; ++m) {
        long double cj[NCOEF];
        for (int j = 0; j < NCOEF; ++j) {
          long double s = 0;
          for (int k = 0; k < NCOEF; ++k)
            s += F[m * NCOEF + k] * cosl(PI * j * (k + 0.5L) / NCOEF);
          cj[j] = s * 2 / NCOEF;
        }
        cj[0] /= 2;
        // p(x) = sum_j cj T_j(x), x = 2 xd  =>  coefficient of xd^i
        for (int i = 0; i < NCOEF; ++i) {
          long double s = 0;
          for (int j = 0; j < NCOEF; ++j) s += cj[j] * tc[j][i];
          c[((size_t)iv * (MMAX + 1) + m) * NCOEF + i] = (double)(s * powl(2.0L, i));
        }
      }
    }
  }
  // restates FmEval_Chebyshev7::eval (boys.h:345-454), scalar path
  void eval(double* Fm, double x, int mmax) const {
    if (x > TCRIT) {
      const double one_over_x = 1 / x;
      Fm[0] = 0.88622692545275801365 * std::sqrt(one_over_x);
      for (int i = 1; i <= mmax; i++) Fm[i] = Fm[i - 1] * (i - 0.5) * one_over_x;
      return;
    }
    const double xod = x * 7.0;
    const int iv = int(xod);
    const double xd = xod - (double)iv - 0.5;
    const double* d = &c[((size_t)iv * (MMAX + 1)) * NCOEF];
    for (int m = 0; m <= mmax; ++m, d += NCOEF)
      Fm[m] = d[0] + xd * (d[1] + xd * (d[2] + xd * (d[3] + xd * (d[4] + xd * (d[5] + xd * (d[6] + xd * d[7]))))));
  }
};
static const Cheb7& cheb7() {
  static Cheb7 inst;
  return inst;
}

// ----------------------------------------------------------------------------
// solid harmonics (solidharmonics.h:114-180), ordering switch (:183-240)
// ----------------------------------------------------------------------------
static int parity(int i) { return (i % 2) ? -1 : 1; }
static double sh_coeff(int l, int m, int lx, int ly, int lz) {
  const int abs_m = std::abs(m);
  if ((lx + ly - abs_m) % 2) return 0.0;
  const int j = (lx + ly - abs_m) / 2;
  if (j < 0) return 0.0;
  const int comp = (m >= 0) ? 1 : -1;
  const int i = abs_m - lx;
  if (comp != parity(std::abs(i))) return 0.0;
  double pfac = std::sqrt(((fac(2 * lx) * fac(2 * ly) * fac(2 * lz)) / fac(2 * l)) *
                          (fac(l - abs_m) / fac(l)) * (1.0 / fac(l + abs_m)) *
                          (1.0 / (fac(lx) * fac(ly) * fac(lz))));
  pfac /= (double)(1L << l);
  pfac *= (m < 0) ? parity((i - 1) / 2) : parity(i / 2);
  const int i_min = j, i_max = (l - abs_m) / 2;
  double sum = 0;
  for (int ii = i_min; ii <= i_max; ii++) {
    double pfac1 = binom(l, ii) * binom(ii, j);
    pfac1 *= (parity(ii) * fac(2 * (l - ii))) / fac(l - abs_m - 2 * ii);
    double sum1 = 0.0;
    const int k_min = std::max((lx - abs_m) / 2, 0);
    const int k_max = std::min(j, lx / 2);
    for (int k = k_min; k <= k_max; k++)
      if (lx - 2 * k <= abs_m) sum1 += binom(j, k) * binom(abs_m, lx - 2 * k) * parity(k);
    sum += pfac1 * sum1;
  }
  sum *= std::sqrt(df_km1(2 * l) / (df_km1(2 * lx) * df_km1(2 * ly) * df_km1(2 * lz)));
  return (m == 0) ? pfac * sum : M_SQRT2 * pfac * sum;
}
static int g_shg_ordering = 0;  // 0 = Standard (m=-l..l), 1 = Gaussian (0,+1,-1,...)
static std::vector<int> shg_m_list(int l) {
  std::vector<int> ms;
  if (g_shg_ordering == 0) {
    for (int m = -l; m <= l; ++m) ms.push_back(m);
  } else {
    for (int m = 0, k = 0; k != 2 * l + 1; ++k, m = (m > 0 ? -m : 1 - m)) ms.push_back(m);
  }
  return ms;
}
// dense (2l+1) x ncart matrix
static std::vector<double> make_sh_matrix(int l) {
  auto ms = shg_m_list(l);
  auto cc = cart_comps(l);
  std::vector<double> M((2 * l + 1) * cc.size());
  for (size_t p = 0; p < ms.size(); ++p)
    for (size_t c = 0; c < cc.size(); ++c) M[p * cc.size() + c] = sh_coeff(l, ms[p], cc[c][0], cc[c][1], cc[c][2]);
  return M;
}
// the matrix for the current ordering, built once per (ordering, l)
static const std::vector<double>& sh_matrix(int l) {
  static std::mutex mtx;
  static std::map<int, std::vector<double>> cache;
  const int key = g_shg_ordering * 64 + l;
  std::lock_guard<std::mutex> lk(mtx);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, make_sh_matrix(l)).first;
  return it->second;
}

// ----------------------------------------------------------------------------
// shells
// ----------------------------------------------------------------------------
struct Shell {
  int l = 0;
  bool pure = false;
  std::vector<double> alpha, coeff, max_ln_coeff;
  std::array<double, 3> O{0, 0, 0};
  int nprim() const { return (int)alpha.size(); }
  int size() const { return nsize(l, pure); }
  int cartesian_size() const { return ncart(l); }
};

// Shell::renorm (shell.h:958-999) + update_max_ln_coeff (:1001-1011)
static void renorm(int l, int np, const double* alpha, double* coeff, double* max_ln_coeff,
                   bool enforce_unit) {
  const double sqrt_Pi_cubed = 5.56832799683170784528481798212;
  for (int p = 0; p < np; ++p) {
    if (alpha[p] != 0) {
      const double two_alpha = 2 * alpha[p];
      const double two_alpha_to_am32 = std::pow(two_alpha, l + 1) * std::sqrt(two_alpha);
      const double nf = std::sqrt(std::pow(2, l) * two_alpha_to_am32 / (sqrt_Pi_cubed * df_km1(2 * l)));
      coeff[p] *= nf;
    }
  }
  if (enforce_unit) {
    double norm = 0;
    for (int p = 0; p < np; ++p)
      for (int q = 0; q <= p; ++q) {
        const double gamma = alpha[p] + alpha[q];
        norm += (p == q ? 1 : 2) * df_km1(2 * l) * sqrt_Pi_cubed * coeff[p] * coeff[q] /
                (std::pow(2, l) * std::pow(gamma, l + 1) * std::sqrt(gamma));
      }
    const double f = 1 / std::sqrt(norm);
    for (int p = 0; p < np; ++p) coeff[p] *= f;
  }
  for (int p = 0; p < np; ++p) max_ln_coeff[p] = std::log(std::abs(coeff[p]));
}

// ----------------------------------------------------------------------------
// shell pairs (shell.h:1137-1256)
// ----------------------------------------------------------------------------
enum Screening { Original = 0x0001, Conservative = 0x0010, Schwarz = 0x0100, SchwarzInf = 0x1000 };

struct PrimPair {
  double P[3], K, one_over_gamma, nonsph, ln_scr;
  int p1, p2;
};
struct ShellPair {
  std::vector<PrimPair> pp;
  double AB[3];
  // Schwarz / SchwarzInf variant (shell.h:1258-1328): factor[p1*n2+p2] is the
  // primitive-pair Schwarz factor supplied by the caller
  void init_schwarz(const Shell& s1, const Shell& s2, double ln_prec, const double* factor) {
    pp.clear();
    const auto& A = s1.O;
    const auto& B = s2.O;
    double AB2 = 0;
    for (int i = 0; i < 3; ++i) {
      AB[i] = A[i] - B[i];
      AB2 += AB[i] * AB[i];
    }
    const int n1 = s1.nprim(), n2 = s2.nprim();
    for (int p1 = 0; p1 < n1; ++p1)
      for (int p2 = 0; p2 < n2; ++p2) {
        const double ln_scr =
            std::log((double)(n1 * n2) * factor[p1 * n2 + p2]) + s1.max_ln_coeff[p1] + s2.max_ln_coeff[p2];
        if (ln_scr < ln_prec) continue;
        const double a1 = s1.alpha[p1], a2 = s2.alpha[p2];
        const double gamma = a1 + a2, oogamma = 1 / gamma;
        const double rho = a1 * a2 * oogamma;
        PrimPair q;
        if (AB2 == 0.) {
          for (int i = 0; i < 3; ++i) q.P[i] = A[i];
        } else {
          for (int i = 0; i < 3; ++i) q.P[i] = (a1 * A[i] + a2 * B[i]) * oogamma;
        }
        q.ln_scr = ln_scr;
        q.p1 = p1;
        q.p2 = p2;
        q.K = 5.9149671727956128778 * std::exp(-rho * AB2) * oogamma;
        q.nonsph = 0;
        q.one_over_gamma = oogamma;
        pp.push_back(q);
      }
  }
  void init(const Shell& s1, const Shell& s2, double ln_prec, int method) {
    pp.clear();
    const auto& A = s1.O;
    const auto& B = s2.O;
    double AB2 = 0;
    for (int i = 0; i < 3; ++i) {
      AB[i] = A[i] - B[i];
      AB2 += AB[i] * AB[i];
    }
    const int n1 = s1.nprim(), n2 = s2.nprim();
    for (int p1 = 0; p1 < n1; ++p1)
      for (int p2 = 0; p2 < n2; ++p2) {
        const double a1 = s1.alpha[p1], a2 = s2.alpha[p2];
        const double gamma = a1 + a2, oogamma = 1 / gamma;
        const double rho = a1 * a2 * oogamma;
        const double mrAB2 = -rho * AB2;
        double ln_scr = mrAB2 + s1.max_ln_coeff[p1] + s2.max_ln_coeff[p2];
        if (method == Original && ln_scr < ln_prec) continue;
        double P[3];
        if (AB2 == 0.) {
          P[0] = A[0], P[1] = A[1], P[2] = A[2];
        } else {
          for (int i = 0; i < 3; ++i) P[i] = (a1 * A[i] + a2 * B[i]) * oogamma;
        }
        double nonsph = 0;
        if (method == Conservative) {
          const double mpa = std::max(std::max(std::abs(P[0] - A[0]), std::abs(P[1] - A[1])), std::abs(P[2] - A[2]));
          const double mpb = std::max(std::max(std::abs(P[0] - B[0]), std::abs(P[1] - B[1])), std::abs(P[2] - B[2]));
          const double t1 = std::pow(mpa, s1.l) * std::pow(mpb, s2.l);
          const double t2 = fac(s1.l) * fac(s2.l) * std::pow(oogamma, s1.l + s2.l);
          nonsph = std::max(t1, t2);
          const double ln_nonsph = std::log(std::max(nonsph, 1.0));
          const double ln_sph_extra = 1.777485947591722872387900 + std::log(oogamma);
          ln_scr += ln_sph_extra + ln_nonsph + std::log((double)(n1 * n2));
          if (ln_scr < ln_prec) continue;
        }
        PrimPair q;
        q.ln_scr = ln_scr;
        q.p1 = p1;
        q.p2 = p2;
        q.K = 5.9149671727956128778 * std::exp(mrAB2) * oogamma;
        for (int i = 0; i < 3; ++i) q.P[i] = P[i];
        q.nonsph = nonsph;
        q.one_over_gamma = oogamma;
        pp.push_back(q);
      }
  }
};

// ----------------------------------------------------------------------------
// Obara-Saika VRR "program" for a canonical class, run vectorised over
// primitive quartets.  Nodes (e, f, m): e = bra Cartesian triple (centre A),
// f = ket Cartesian triple (centre C), m = auxiliary index.  The recursion
// (vrr_11_twoprep_11.h:152-224) used here:
//  bra:  [e+1i|0]^m = PA_i [e|0]^m + WP_i [e|0]^{m+1}
//                     + e_i/(2z) ([e-1i|0]^m - rho/z [e-1i|0]^{m+1})
//  ket:  [e|f+1i]^m = QC_i [e|f]^m + WQ_i [e|f]^{m+1}
//                     + f_i/(2n) ([e|f-1i]^m - rho/n [e|f-1i]^{m+1})
//                     + e_i/(2(z+n)) [e-1i|f]^{m+1}
// The build direction is the first axis (x, y, z order) with a nonzero
// quantum number.
// ----------------------------------------------------------------------------
struct VrrOp {
  int dst, kind;  // kind 0 = base [0|0]^m, 1 = bra step, 2 = ket step
  int dir, m;
  int s0, s1, s2, s3, s4;  // operand slots (-1 = absent)
  double c23, c4;          // integer multiplicities e_i / f_i, e_i (ket cross)
};
struct VrrProgram {
  int la, lb, lc, ld, L;
  int nslots;
  std::vector<VrrOp> ops;
  // target (e index in [la..la+lb] block, f index in [lc..lc+ld] block) -> slot
  int ne, nf;
  std::vector<int> target_slot;  // ne*nf
};

struct Key {
  int e[3], f[3], m;
  bool operator<(const Key& o) const {
    return std::memcmp(this, &o, sizeof(Key)) < 0;
  }
};

static int first_dir(const int* v) {
  for (int i = 0; i < 3; ++i)
    if (v[i] > 0) return i;
  return -1;
}

static VrrProgram make_vrr_program(int la, int lb, int lc, int ld) {
  VrrProgram P;
  P.la = la, P.lb = lb, P.lc = lc, P.ld = ld;
  P.L = la + lb + lc + ld;
  std::vector<std::array<int, 3>> ecomps, fcomps;
  for (int l = la; l <= la + lb; ++l)
    for (auto& c : cart_comps(l)) ecomps.push_back(c);
  for (int l = lc; l <= lc + ld; ++l)
    for (auto& c : cart_comps(l)) fcomps.push_back(c);
  P.ne = (int)ecomps.size();
  P.nf = (int)fcomps.size();
  std::map<Key, int> node_id;
  std::vector<Key> nodes;
  std::vector<std::vector<int>> deps;
  // recursive construction (post-order = topological)
  std::function<int(const Key&)> visit;
  std::vector<VrrOp> proto;
  visit = [&](const Key& k) -> int {
    auto it = node_id.find(k);
    if (it != node_id.end()) return it->second;
    VrrOp op{};
    op.m = k.m;
    op.s0 = op.s1 = op.s2 = op.s3 = op.s4 = -1;
    op.c23 = op.c4 = 0;
    const int fsum = k.f[0] + k.f[1] + k.f[2];
    const int esum = k.e[0] + k.e[1] + k.e[2];
    if (fsum > 0) {
      const int i = first_dir(k.f);
      Key a = k;
      a.f[i] -= 1;
      Key b = a;
      b.m += 1;
      op.kind = 2;
      op.dir = i;
      op.s0 = visit(a);
      op.s1 = visit(b);
      if (a.f[i] > 0) {
        Key c = a;
        c.f[i] -= 1;
        Key d = c;
        d.m += 1;
        op.s2 = visit(c);
        op.s3 = visit(d);
        op.c23 = a.f[i];
      }
      if (k.e[i] > 0) {
        Key c = a;
        c.e[i] -= 1;
        c.m += 1;
        op.s4 = visit(c);
        op.c4 = k.e[i];
      }
    } else if (esum > 0) {
      const int i = first_dir(k.e);
      Key a = k;
      a.e[i] -= 1;
      Key b = a;
      b.m += 1;
      op.kind = 1;
      op.dir = i;
      op.s0 = visit(a);
      op.s1 = visit(b);
      if (a.e[i] > 0) {
        Key c = a;
        c.e[i] -= 1;
        Key d = c;
        d.m += 1;
        op.s2 = visit(c);
        op.s3 = visit(d);
        op.c23 = a.e[i];
      }
    } else {
      op.kind = 0;
      op.dir = -1;
    }
    const int id = (int)nodes.size();
    nodes.push_back(k);
    op.dst = id;
    proto.push_back(op);
    node_id[k] = id;
    return id;
  };
  P.target_slot.resize(P.ne * P.nf);
  for (int ie = 0; ie < P.ne; ++ie)
    for (int jf = 0; jf < P.nf; ++jf) {
      Key k{};
      for (int i = 0; i < 3; ++i) k.e[i] = ecomps[ie][i], k.f[i] = fcomps[jf][i];
      k.m = 0;
      P.target_slot[ie * P.nf + jf] = visit(k);
    }
  P.ops = proto;
  P.nslots = (int)nodes.size();
  return P;
}

static const VrrProgram& vrr_program(int la, int lb, int lc, int ld) {
  static std::mutex mtx;
  static std::map<int, std::unique_ptr<VrrProgram>> cache;
  const int key = ((la * 16 + lb) * 16 + lc) * 16 + ld;
  std::lock_guard<std::mutex> lk(mtx);
  auto it = cache.find(key);
  if (it != cache.end()) return *it->second;
  auto p = std::make_unique<VrrProgram>(make_vrr_program(la, lb, lc, ld));
  auto& r = *p;
  cache[key] = std::move(p);
  return r;
}

// ----------------------------------------------------------------------------
// HRR on one index pair of a contracted array: transfers from
// X[e in (l1..l1+l2) block][rest] to Y[a in l1][b in l2][rest] with
// (a, b+1i) = (a+1i, b) + AB_i (a, b)   (hrr.h:290-330).  Generic, level by
// level in b; the build direction is the first axis with b_i > 0.
// ----------------------------------------------------------------------------
// The transfer runs from a program built once per (l1, l2): per level bl of
// b, the size of the level's array and its ops (destination, the two
// sources as row offsets into the previous level, the axis).
struct HrrOp {
  int d, s0, s1, dir;
};
struct HrrProgram {
  int l1, l2;
  int nin;                           // rows of X: the e blocks l1 .. l1+l2
  std::vector<int> nrows;            // rows of level bl = 1 .. l2
  std::vector<std::vector<HrrOp>> ops;
};
static HrrProgram make_hrr_program(int l1, int l2) {
  HrrProgram P;
  P.l1 = l1, P.l2 = l2;
  P.nin = 0;
  for (int l = l1; l <= l1 + l2; ++l) P.nin += ncart(l);
  for (int bl = 1; bl <= l2; ++bl) {
    const int amax = l1 + l2 - bl;
    // rows of the level: (a of level al in l1 .. amax, index ia) x b index
    std::vector<int> aoff(amax - l1 + 2, 0), aoffp(amax + 1 - l1 + 2, 0);
    for (int l = l1; l <= amax; ++l) aoff[l + 1 - l1] = aoff[l - l1] + ncart(l);
    for (int l = l1; l <= amax + 1; ++l) aoffp[l + 1 - l1] = aoffp[l - l1] + ncart(l);
    const int nb = ncart(bl), nbm = ncart(bl - 1);
    const auto bcomps = cart_comps(bl);
    std::vector<HrrOp> lev;
    for (int al = l1; al <= amax; ++al) {
      const auto acomps = cart_comps(al);
      for (size_t ia = 0; ia < acomps.size(); ++ia)
        for (int ib = 0; ib < nb; ++ib) {
          const auto& b = bcomps[ib];
          const int i = first_dir(b.data());
          std::array<int, 3> bm = b;
          bm[i] -= 1;
          const int ibm = cart_index(bm[0], bm[1], bm[2]);
          std::array<int, 3> ap = acomps[ia];
          ap[i] += 1;
          const int iap = cart_index(ap[0], ap[1], ap[2]);
          HrrOp op;
          op.d = (aoff[al - l1] + (int)ia) * nb + ib;
          op.s0 = (aoffp[al + 1 - l1] + iap) * nbm + ibm;
          op.s1 = (aoffp[al - l1] + (int)ia) * nbm + ibm;
          op.dir = i;
          lev.push_back(op);
        }
    }
    P.nrows.push_back(aoff[amax - l1 + 1] * nb);
    P.ops.push_back(std::move(lev));
  }
  return P;
}
static const HrrProgram& hrr_program(int l1, int l2) {
  static std::mutex mtx;
  static std::map<int, std::unique_ptr<HrrProgram>> cache;
  const int key = l1 * 64 + l2;
  std::lock_guard<std::mutex> lk(mtx);
  auto it = cache.find(key);
  if (it != cache.end()) return *it->second;
  auto p = std::make_unique<HrrProgram>(make_hrr_program(l1, l2));
  auto& r = *p;
  cache[key] = std::move(p);
  return r;
}
// cur and nxt: caller-owned work arrays (reused across calls)
template <class Real>
static void hrr_transfer(int l1, int l2, const Real AB[3], const Real* X, int rest, Real* Y, std::vector<Real>& cur,
                         std::vector<Real>& nxt) {
  if (l2 == 0) {
    std::memcpy(Y, X, sizeof(Real) * ncart(l1) * rest);
    return;
  }
  const HrrProgram& P = hrr_program(l1, l2);
  const Real* src = X;
  for (int bl = 1; bl <= l2; ++bl) {
    std::vector<Real>& dstv = (bl & 1) ? cur : nxt;
    dstv.resize((size_t)P.nrows[bl - 1] * rest);
    Real* dst = dstv.data();
    for (const HrrOp& op : P.ops[bl - 1]) {
      const Real* s0 = src + (size_t)op.s0 * rest;
      const Real* s1 = src + (size_t)op.s1 * rest;
      Real* d = dst + (size_t)op.d * rest;
      const Real ab = AB[op.dir];
      for (int r = 0; r < rest; ++r) d[r] = s0[r] + ab * s1[r];
    }
    src = dst;
  }
  std::memcpy(Y, src, sizeof(Real) * ncart(l1) * ncart(l2) * rest);
}

// Core integrals of the operator (Engine's core_eval_pack, engine.impl.h:
// 1437-1479): Coulomb F_m(T); erf_coulomb and erfc_coulomb as
// erf_coulomb_gm_eval / erfc_coulomb_gm_eval (boys.h:1913-1985):
//   erf:  G_m = s^(m+1/2) F_m(s T),  s = omega^2 / (omega^2 + rho)  (0 if omega = 0)
//   erfc: G_m = F_m(T) - s^(m+1/2) F_m(s T)
// HRR build centre of an equal-l pair: the reference keeps the caller's
// order (engine.impl.h:1181-1208).  With g_tight_centre the restatement
// builds an f pair instead on the shell the pair's charge sits nearer (a
// K-weighted comparison of |P - A| and |P - B|), as the GPU library does:
// the same integrals, with far less HRR cancellation for a distant pair of
// a diffuse and a tight shell.  Never used by the reference restatement.
static int g_tight_centre = 0;
static int g_progress = 0;  // fock_build prints its progress to stderr (long runs)
// fock_build digests only the quartets with at least one shell in g_rows
// (when not empty): every quartet that adds to a row or column of those
// shells' functions, so their rows of G, J and K come out whole (the
// sampled-row check of large Fock builds)
static std::vector<int8_t> g_rows;
static int g_oper = 0;  // 0 coulomb, 1 erf_coulomb, 2 erfc_coulomb
static double g_omega = 0;
template <class Boys>
static void core_ints(const Boys& boys, double* Gm, double rho, double T, int mmax) {
  if (g_oper == 0) {
    boys.eval(Gm, T, mmax);
    return;
  }
  if (g_oper == 2)
    boys.eval(Gm, T, mmax);
  else
    for (int m = 0; m <= mmax; ++m) Gm[m] = 0.0;
  if (g_omega > 0) {
    const double o2 = g_omega * g_omega, s = o2 / (o2 + rho);
    double Fs[64];
    boys.eval(Fs, T * s, mmax);
    double f = std::sqrt(s);
    for (int m = 0; m <= mmax; ++m, f *= s) Gm[m] += (g_oper == 2 ? -f : f) * Fs[m];
  }
}

// The same core integrals in long double for the extended-precision
// restatement (eri_quartets_ld), from FmEval_Reference2's series.
static void core_ints_ld(long double* Gm, long double rho, long double T, int mmax) {
  if (g_oper == 0 || g_oper == 2)
    boys_reference(T, mmax, Gm);
  else
    for (int m = 0; m <= mmax; ++m) Gm[m] = 0;
  if (g_omega > 0) {
    const long double o2 = (long double)g_omega * g_omega, s = o2 / (o2 + rho);
    long double Fs[64];
    boys_reference(T * s, mmax, Fs);
    long double f = sqrtl(s);
    for (int m = 0; m <= mmax; ++m, f *= s) Gm[m] += (g_oper == 2 ? -f : f) * Fs[m];
  }
}
// Quad precision (__float128, 113-bit significand): the Boys function by
// its series below 117 and erf + upward recursion above, the attenuated
// forms as in long double.  For the arbiter of values the double and long
// double restatements both lose to cancellation.
typedef __float128 quad;
static quad boys_series_q(quad T, int m) {
  quad denom = m + quad(0.5);
  quad term = expq(-T) / (2 * denom);
  quad sum = term, old;
  const quad eps = FLT128_EPSILON / 10;
  do {
    denom += 1;
    old = term;
    term = old * T / denom;
    sum += term;
  } while (term > sum * eps || old < term);
  return sum;
}
static void boys_reference_q(quad T, int mmax, quad* F) {
  if (T < 117) {
    for (int m = 0; m <= mmax; ++m) F[m] = boys_series_q(T, m);
  } else {
    const quad t2 = 2 * T, et = expq(-T), st = sqrtq(T);
    F[0] = sqrtq(M_PIq) / 2 * erfq(st) / st;
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - et) / t2;
  }
}
static inline void core_ints_t(quad* Gm, quad rho, quad T, int mmax) {
  if (g_oper == 0 || g_oper == 2)
    boys_reference_q(T, mmax, Gm);
  else
    for (int m = 0; m <= mmax; ++m) Gm[m] = 0;
  if (g_omega > 0) {
    const quad o2 = (quad)g_omega * g_omega, s = o2 / (o2 + rho);
    quad Fs[64];
    boys_reference_q(T * s, mmax, Fs);
    quad f = sqrtq(s);
    for (int m = 0; m <= mmax; ++m, f *= s) Gm[m] += (g_oper == 2 ? -f : f) * Fs[m];
  }
}
static inline quad sqrt(quad x) { return sqrtq(x); }
static inline void core_ints_t(double* Gm, double rho, double T, int mmax) { core_ints(cheb7(), Gm, rho, T, mmax); }
static inline void core_ints_t(long double* Gm, long double rho, long double T, int mmax) {
  core_ints_ld(Gm, rho, T, mmax);
}

// Primitive-pair quantities.  In double they are the ShellPair's stored
// values, exactly as the reference reads them (engine.impl.h:1330-1360); in
// long double they are recomputed from the shells in long double, so that
// the extended-precision restatement carries no double rounding of its
// inputs beyond the exponents, coefficients and centres themselves.
template <class Real>
static inline Real oo_gamma(double a1, double a2, double stored) {
  if constexpr (std::is_same_v<Real, double>) return stored;
  return Real(1) / ((Real)a1 + (Real)a2);
}
template <class Real>
static inline void pair_center(const struct Shell& s1, const struct Shell& s2, int p1, int p2, const double* stored,
                               Real* P);
template <class Real>
static inline Real pair_K(const struct Shell& s1, const struct Shell& s2, int p1, int p2, double stored);

// ----------------------------------------------------------------------------
// Engine restatement
// ----------------------------------------------------------------------------
template <class Real>
static inline void pair_center(const Shell& s1, const Shell& s2, int p1, int p2, const double* stored, Real* P) {
  if constexpr (std::is_same_v<Real, double>) {
    for (int i = 0; i < 3; ++i) P[i] = stored[i];
  } else {
    const Real a1 = s1.alpha[p1], a2 = s2.alpha[p2];
    bool same = true;
    for (int i = 0; i < 3; ++i) same = same && s1.O[i] == s2.O[i];
    for (int i = 0; i < 3; ++i) P[i] = same ? (Real)s1.O[i] : (a1 * (Real)s1.O[i] + a2 * (Real)s2.O[i]) / (a1 + a2);
  }
}
template <class Real>
static inline Real pair_K(const Shell& s1, const Shell& s2, int p1, int p2, double stored) {
  if constexpr (std::is_same_v<Real, double>) {
    return stored;
  } else {
    const Real a1 = s1.alpha[p1], a2 = s2.alpha[p2], g = a1 + a2;
    Real AB2 = 0;
    for (int i = 0; i < 3; ++i) AB2 += ((Real)s1.O[i] - (Real)s2.O[i]) * ((Real)s1.O[i] - (Real)s2.O[i]);
    // sqrt(2) pi^(5/4) exp(-a1 a2 / g |AB|^2) / g (shell.h:1244)
    if constexpr (std::is_same_v<Real, quad>)
      return sqrtq(2 * powq(M_PIq, quad(2.5))) * expq(-a1 * a2 / g * AB2) / g;
    else
      return 5.914967172795612877839928522935L * expl(-a1 * a2 / g * AB2) / g;
  }
}

// true when the second shell of an equal-l pair is the better build centre
static bool second_is_centre(const Shell& a, const Shell& b) {
  if (a.l != b.l) return false;
  double AB2 = 0;
  for (int k = 0; k < 3; ++k) AB2 += (a.O[k] - b.O[k]) * (a.O[k] - b.O[k]);
  if (AB2 == 0) return false;
  // shells below f keep the caller's order: with both shells of a pair
  // mixing tight and diffuse primitives neither centre is reliably better,
  // and on d pairs of the s4p3d2 recipe the caller's order measured closer
  // to the exact integrals; on f pairs the tighter shell cuts the f-recipe
  // Fock error against the exact one from 2e-6 to 3e-11
  if (a.l < 3) return false;
  double s = 0;
  for (int p1 = 0; p1 < a.nprim(); ++p1)
    for (int p2 = 0; p2 < b.nprim(); ++p2) {
      const double a1 = a.alpha[p1], a2 = b.alpha[p2], g = a1 + a2;
      const double w = std::fabs(a.coeff[p1] * b.coeff[p2]) * std::exp(-a1 * a2 / g * AB2) / g;
      s += w * (a2 - a1) / g;
    }
  return s > 0;
}

template <class Real>
struct EngineT {
  double precision = std::numeric_limits<double>::epsilon();
  double ln_precision = std::log(std::numeric_limits<double>::epsilon());
  int screening = Original;
  double scale = 1.0;
  bool uniform_cart = false;
  int tight = -1;  // HRR build centre of equal-l pairs: -1 as g_tight_centre, 0 caller's order, 1 tighter shell
  // scratch
  ShellPair spbra, spket;
  std::vector<Real> prim[16];
  std::vector<Real> buf, acc, work1, work2;
  std::vector<Real> roe, gm, cart, kt, cur, nx, h1, h2;  // reused: no allocation once warm
  std::vector<Real> result;  // caller-order result
  void set_precision(double p) {
    if (p <= 0) {
      precision = 0;
      ln_precision = std::numeric_limits<double>::lowest();
    } else {
      precision = p;
      ln_precision = std::log(p);
    }
  }

  // returns false if screened out entirely (reference returns nullptr)
  bool compute2(const Shell& tb1, const Shell& tb2, const Shell& tk1, const Shell& tk2,
                const ShellPair* tspbra = nullptr, const ShellPair* tspket = nullptr);
};

template <class Real>
bool EngineT<Real>::compute2(const Shell& tb1, const Shell& tb2, const Shell& tk1, const Shell& tk2,
                      const ShellPair* tspbra, const ShellPair* tspket) {
  // canonical reorder (engine.impl.h:1181-1208)
  const bool tc = tight >= 0 ? tight != 0 : g_tight_centre != 0;
  const bool swap_tbra = tb1.l < tb2.l || (tc && second_is_centre(tb1, tb2));
  const bool swap_tket = tk1.l < tk2.l || (tc && second_is_centre(tk1, tk2));
  const bool swap_braket = (tb1.l + tb2.l) > (tk1.l + tk2.l);
  const Shell& b1 = swap_braket ? (swap_tket ? tk2 : tk1) : (swap_tbra ? tb2 : tb1);
  const Shell& b2 = swap_braket ? (swap_tket ? tk1 : tk2) : (swap_tbra ? tb1 : tb2);
  const Shell& k1 = swap_braket ? (swap_tbra ? tb2 : tb1) : (swap_tket ? tk2 : tk1);
  const Shell& k2 = swap_braket ? (swap_tbra ? tb1 : tb2) : (swap_tket ? tk1 : tk2);
  const bool swap_bra = swap_braket ? swap_tket : swap_tbra;
  const bool swap_ket = swap_braket ? swap_tbra : swap_tket;
  const ShellPair* spb_pre = swap_braket ? tspket : tspbra;
  const ShellPair* spk_pre = swap_braket ? tspbra : tspket;

  const ShellPair* spb;
  const ShellPair* spk;
  bool spb_swapped = false, spk_swapped = false;
  if (spb_pre) {
    spb = spb_pre;
    spb_swapped = swap_bra;
  } else {
    spbra.init(b1, b2, ln_precision, screening);
    spb = &spbra;
  }
  if (spk_pre) {
    spk = spk_pre;
    spk_swapped = swap_ket;
  } else {
    spket.init(k1, k2, ln_precision, screening);
    spk = &spket;
  }
  Real AB[3], CD[3];
  for (int i = 0; i < 3; ++i) {
    if constexpr (std::is_same_v<Real, double>) {
      AB[i] = spb_swapped ? -spb->AB[i] : spb->AB[i];
      CD[i] = spk_swapped ? -spk->AB[i] : spk->AB[i];
    } else {
      AB[i] = (Real)b1.O[i] - (Real)b2.O[i];
      CD[i] = (Real)k1.O[i] - (Real)k2.O[i];
    }
  }
  const auto& A = b1.O;
  const auto& C = k1.O;
  const int la = b1.l, lb = b2.l, lc = k1.l, ld = k2.l;
  const int L = la + lb + lc + ld;
  const int npb = (int)spb->pp.size(), npk = (int)spk->pp.size();
  const double npbraket = (double)npb * npk;
  const int maxp = npb * npk;
  // per-primitive-quartet prefactors, SoA
  enum { PAx, PAy, PAz, WPx, WPy, WPz, QCx, QCy, QCz, WQx, WQy, WQz, OO2Z, OO2E, OO2ZE, ROZ };
  for (auto& v : prim) v.resize(maxp);
  roe.resize(maxp);
  gm.resize((size_t)(L + 1) * maxp);
  Real Fm[64];
  int p = 0;
  for (int ib = 0; ib < npb; ++ib)
    for (int ik = 0; ik < npk; ++ik) {
      const PrimPair& bp = spb->pp[ib];
      const PrimPair& kp = spk->pp[ik];
      if (!(bp.ln_scr + kp.ln_scr > ln_precision)) continue;
      const int pb1 = spb_swapped ? bp.p2 : bp.p1, pb2 = spb_swapped ? bp.p1 : bp.p2;
      const int pk1 = spk_swapped ? kp.p2 : kp.p1, pk2 = spk_swapped ? kp.p1 : kp.p2;
      const Real a0 = b1.alpha[pb1], a1 = b2.alpha[pb2], a2 = k1.alpha[pk1], a3 = k2.alpha[pk2];
      const Real c0 = b1.coeff[pb1], c1 = b2.coeff[pb2], c2 = k1.coeff[pk1], c3 = k2.coeff[pk2];
      const Real gp = a0 + a1, oogp = oo_gamma<Real>(a0, a1, bp.one_over_gamma);
      const Real gq = a2 + a3, oogq = oo_gamma<Real>(a2, a3, kp.one_over_gamma);
      Real P[3], Q[3];
      pair_center<Real>(b1, b2, pb1, pb2, bp.P, P);
      pair_center<Real>(k1, k2, pk1, pk2, kp.P, Q);
      const Real PQx = P[0] - Q[0], PQy = P[1] - Q[1], PQz = P[2] - Q[2];
      const Real PQ2 = PQx * PQx + PQy * PQy + PQz * PQz;
      const Real K12 = pair_K<Real>(b1, b2, pb1, pb2, bp.K) * pair_K<Real>(k1, k2, pk1, pk2, kp.K);
      const Real gpq = gp + gq;
      const Real sqrt_gpq = sqrt(gpq);
      const Real oogpq = Real(1) / gpq;
      Real pfac = K12 * sqrt_gpq * oogpq;
      pfac *= c0 * c1 * c2 * c3 * Real(scale);
      if (screening & (Original | Conservative)) {
        double mag = std::abs((double)pfac);
        if (screening == Conservative) mag *= std::max(1., bp.nonsph * kp.nonsph) * npbraket;
        if (mag < precision) continue;
      }
      const Real rho = gp * gq * oogpq;
      const Real T = PQ2 * rho;
      core_ints_t(Fm, rho, T, L);
      for (int m = 0; m <= L; ++m) gm[(size_t)m * maxp + p] = Fm[m] * pfac;
      const Real gpo = oogpq * gp, gqo = oogpq * gq;
      Real W[3];
      for (int i = 0; i < 3; ++i) W[i] = gpo * P[i] + gqo * Q[i];
      for (int i = 0; i < 3; ++i) {
        prim[PAx + i][p] = P[i] - A[i];
        prim[WPx + i][p] = W[i] - P[i];
        prim[QCx + i][p] = Q[i] - C[i];
        prim[WQx + i][p] = W[i] - Q[i];
      }
      prim[OO2Z][p] = 0.5 * oogp;
      prim[OO2E][p] = 0.5 * oogq;
      prim[OO2ZE][p] = 0.5 * oogpq;
      prim[ROZ][p] = rho * oogp;
      roe[p] = rho * oogq;
      ++p;
    }
  const int np = p;
  if (np == 0) return false;

  const int n1 = b1.size(), n2 = b2.size(), n3 = k1.size(), n4 = k2.size();
  const int c1n = ncart(la), c2n = ncart(lb), c3n = ncart(lc), c4n = ncart(ld);
  cart.resize((size_t)c1n * c2n * c3n * c4n);
  if (L == 0) {
    Real s = 0;
    for (int q = 0; q < np; ++q) s += gm[q];
    cart[0] = s;
  } else {
    const VrrProgram& prog = vrr_program(la, lb, lc, ld);
    buf.resize((size_t)prog.nslots * np);
    for (const auto& op : prog.ops) {
      Real* d = &buf[(size_t)op.dst * np];
      if (op.kind == 0) {
        const Real* g = &gm[(size_t)op.m * maxp];
        for (int q = 0; q < np; ++q) d[q] = g[q];
        continue;
      }
      const Real* x0 = &buf[(size_t)op.s0 * np];
      const Real* x1 = &buf[(size_t)op.s1 * np];
      if (op.kind == 1) {
        const Real* pa = prim[PAx + op.dir].data();
        const Real* wp = prim[WPx + op.dir].data();
        if (op.s2 >= 0) {
          const Real* x2 = &buf[(size_t)op.s2 * np];
          const Real* x3 = &buf[(size_t)op.s3 * np];
          const Real* z = prim[OO2Z].data();
          const Real* r = prim[ROZ].data();
          const Real c = op.c23;
          for (int q = 0; q < np; ++q)
            d[q] = pa[q] * x0[q] + wp[q] * x1[q] + c * z[q] * (x2[q] - r[q] * x3[q]);
        } else {
          for (int q = 0; q < np; ++q) d[q] = pa[q] * x0[q] + wp[q] * x1[q];
        }
      } else {
        const Real* qc = prim[QCx + op.dir].data();
        const Real* wq = prim[WQx + op.dir].data();
        for (int q = 0; q < np; ++q) d[q] = qc[q] * x0[q] + wq[q] * x1[q];
        if (op.s2 >= 0) {
          const Real* x2 = &buf[(size_t)op.s2 * np];
          const Real* x3 = &buf[(size_t)op.s3 * np];
          const Real* e = prim[OO2E].data();
          const Real c = op.c23;
          for (int q = 0; q < np; ++q) d[q] += c * e[q] * (x2[q] - roe[q] * x3[q]);
        }
        if (op.s4 >= 0) {
          const Real* x4 = &buf[(size_t)op.s4 * np];
          const Real* ze = prim[OO2ZE].data();
          const Real c = op.c4;
          for (int q = 0; q < np; ++q) d[q] += c * ze[q] * x4[q];
        }
      }
    }
    // contraction (dg.cc:1177-1189): sum targets over primitive quartets
    const int ne = prog.ne, nf = prog.nf;
    acc.assign((size_t)ne * nf, 0.0);
    for (int t = 0; t < ne * nf; ++t) {
      const Real* x = &buf[(size_t)prog.target_slot[t] * np];
      Real s = 0;
      for (int q = 0; q < np; ++q) s += x[q];
      acc[t] = s;
    }
    // HRR on the bra: X[e][f] -> [a b][f]
    work1.resize((size_t)c1n * c2n * nf);
    hrr_transfer(la, lb, AB, acc.data(), nf, work1.data(), h1, h2);
    // HRR on the ket: for each (ab), [f] -> [c d]
    // transpose so that f is leading: T[f][ab]
    const int nab = c1n * c2n;
    work2.resize((size_t)nf * nab);
    for (int ab = 0; ab < nab; ++ab)
      for (int f = 0; f < nf; ++f) work2[(size_t)f * nab + ab] = work1[(size_t)ab * nf + f];
    kt.resize((size_t)c3n * c4n * nab);
    hrr_transfer(lc, ld, CD, work2.data(), nab, kt.data(), h1, h2);
    for (int cd = 0; cd < c3n * c4n; ++cd)
      for (int ab = 0; ab < nab; ++ab) cart[(size_t)ab * c3n * c4n + cd] = kt[(size_t)cd * nab + ab];
  }
  // cart -> sph per index (solidharmonics transform_first/inner/last)
  cur.assign(cart.begin(), cart.end());
  int dims[4] = {c1n, c2n, c3n, c4n};
  const Shell* sh[4] = {&b1, &b2, &k1, &k2};
  for (int k = 0; k < 4; ++k) {
    if (!sh[k]->pure) continue;
    const int l = sh[k]->l;
    const auto& M = sh_matrix(l);
    const int nc = ncart(l), ns = 2 * l + 1;
    int pre = 1, post = 1;
    for (int j = 0; j < k; ++j) pre *= dims[j];
    for (int j = k + 1; j < 4; ++j) post *= dims[j];
    nx.assign((size_t)pre * ns * post, 0.0);
    for (int a = 0; a < pre; ++a)
      for (int s = 0; s < ns; ++s)
        for (int c = 0; c < nc; ++c) {
          const Real v = M[s * nc + c];
          if (v == 0) continue;
          const Real* src = &cur[((size_t)a * nc + c) * post];
          Real* dst = &nx[((size_t)a * ns + s) * post];
          for (int r = 0; r < post; ++r) dst[r] += v * src[r];
        }
    cur.swap(nx);
    dims[k] = ns;
  }
  // un-permute into caller order (engine.impl.h:1987-2067)
  const int t1 = tb1.size(), t2 = tb2.size(), t3 = tk1.size(), t4 = tk2.size();
  const int ts[4] = {t2 * t3 * t4, t3 * t4, t4, 1};
  // caller position of each canonical index
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
  result.assign((size_t)t1 * t2 * t3 * t4, 0.0);
  for (int i1 = 0; i1 < n1; ++i1)
    for (int i2 = 0; i2 < n2; ++i2)
      for (int i3 = 0; i3 < n3; ++i3)
        for (int i4 = 0; i4 < n4; ++i4) {
          const size_t src = (((size_t)i1 * n2 + i2) * n3 + i3) * n4 + i4;
          const size_t dst = (size_t)i1 * ts[pos[0]] + (size_t)i2 * ts[pos[1]] + (size_t)i3 * ts[pos[2]] +
                             (size_t)i4 * ts[pos[3]];
          result[dst] = cur[src];
        }
  // uniform Cartesian normalization (cartesian.h:117-133)
  if (uniform_cart) {
    const Shell* ts4[4] = {&tb1, &tb2, &tk1, &tk2};
    std::vector<double> sc[4];
    for (int k = 0; k < 4; ++k) {
      const int l = ts4[k]->l;
      if (ts4[k]->pure) {
        sc[k].assign(2 * l + 1, 1.0);
      } else {
        for (auto& c : cart_comps(l))
          sc[k].push_back(std::sqrt(df_km1(2 * l) / (df_km1(2 * c[0]) * df_km1(2 * c[1]) * df_km1(2 * c[2]))));
      }
    }
    size_t idx = 0;
    for (int i1 = 0; i1 < t1; ++i1)
      for (int i2 = 0; i2 < t2; ++i2)
        for (int i3 = 0; i3 < t3; ++i3)
          for (int i4 = 0; i4 < t4; ++i4, ++idx) result[idx] *= sc[0][i1] * sc[1][i2] * sc[2][i3] * sc[3][i4];
  }
  return true;
}

using Engine = EngineT<double>;

// ----------------------------------------------------------------------------
// brute force: McMurchie-Davidson over primitive Cartesian Gaussians in
// long double; contracted with the renormalised coefficients, then the same
// cart->sph coefficients, written directly in caller order.
// ----------------------------------------------------------------------------
static void md_E(int imax, int jmax, long double a, long double b, long double XAB, std::vector<long double>& E) {
  // E[i][j][t], t <= i+j
  const long double p = a + b, mu = a * b / p;
  const long double XPA = -b / p * XAB, XPB = a / p * XAB;
  const int T = imax + jmax + 1;
  E.assign((size_t)(imax + 1) * (jmax + 1) * (T + 1), 0.0L);
  auto at = [&](int i, int j, int t) -> long double& { return E[((size_t)i * (jmax + 1) + j) * (T + 1) + t]; };
  auto get = [&](int i, int j, int t) -> long double {
    if (t < 0 || t > i + j || i < 0 || j < 0) return 0.0L;
    return E[((size_t)i * (jmax + 1) + j) * (T + 1) + t];
  };
  at(0, 0, 0) = expl(-mu * XAB * XAB);
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        if (i > 0)
          at(i, j, t) = get(i - 1, j, t - 1) / (2 * p) + XPA * get(i - 1, j, t) + (t + 1) * get(i - 1, j, t + 1);
        else
          at(i, j, t) = get(i, j - 1, t - 1) / (2 * p) + XPB * get(i, j - 1, t) + (t + 1) * get(i, j - 1, t + 1);
      }
    }
}

static long double md_primitive(const int* a, const int* b, const int* c, const int* d, long double alpha,
                                long double beta, long double gamma, long double delta, const long double* A,
                                const long double* B, const long double* C, const long double* D) {
  const long double p = alpha + beta, q = gamma + delta;
  long double P[3], Q[3];
  for (int i = 0; i < 3; ++i) {
    P[i] = (alpha * A[i] + beta * B[i]) / p;
    Q[i] = (gamma * C[i] + delta * D[i]) / q;
  }
  const long double al = p * q / (p + q);
  const long double XPQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
  const int lab[3] = {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
  const int lcd[3] = {c[0] + d[0], c[1] + d[1], c[2] + d[2]};
  const int Lt = lab[0] + lab[1] + lab[2] + lcd[0] + lcd[1] + lcd[2];
  std::vector<long double> Eb[3], Ek[3];
  for (int i = 0; i < 3; ++i) {
    md_E(a[i], b[i], alpha, beta, A[i] - B[i], Eb[i]);
    md_E(c[i], d[i], gamma, delta, C[i] - D[i], Ek[i]);
  }
  auto Eget = [](const std::vector<long double>& E, int imax, int jmax, int i, int j, int t) {
    const int T = imax + jmax + 1;
    return E[((size_t)i * (jmax + 1) + j) * (T + 1) + t];
  };
  // Hermite sum with the Coulomb exponent al (or the attenuated one)
  auto hsum = [&](long double al) -> long double {
  // R^n_{tuv}
  std::vector<long double> Fb(Lt + 1);
  boys_reference(al * (XPQ[0] * XPQ[0] + XPQ[1] * XPQ[1] + XPQ[2] * XPQ[2]), Lt, Fb.data());
  const int S = Lt + 1;
  auto idx = [S](int n, int t, int u, int v) { return (((size_t)n * S + t) * S + u) * S + v; };
  std::vector<long double> R((size_t)S * S * S * S, 0.0L);
  for (int n = 0; n <= Lt; ++n) R[idx(n, 0, 0, 0)] = powl(-2 * al, n) * Fb[n];
  for (int n = Lt - 1; n >= 0; --n) {
    const int top = Lt - n;
    for (int t = 0; t <= top; ++t)
      for (int u = 0; u <= top - t; ++u)
        for (int v = 0; v <= top - t - u; ++v) {
          if (t + u + v == 0) continue;
          long double val;
          if (t > 0) {
            val = XPQ[0] * R[idx(n + 1, t - 1, u, v)];
            if (t > 1) val += (t - 1) * R[idx(n + 1, t - 2, u, v)];
          } else if (u > 0) {
            val = XPQ[1] * R[idx(n + 1, t, u - 1, v)];
            if (u > 1) val += (u - 1) * R[idx(n + 1, t, u - 2, v)];
          } else {
            val = XPQ[2] * R[idx(n + 1, t, u, v - 1)];
            if (v > 1) val += (v - 1) * R[idx(n + 1, t, u, v - 2)];
          }
          R[idx(n, t, u, v)] = val;
        }
  }
  long double sum = 0;
  for (int t = 0; t <= lab[0]; ++t)
    for (int u = 0; u <= lab[1]; ++u)
      for (int v = 0; v <= lab[2]; ++v) {
        const long double eb = Eget(Eb[0], a[0], b[0], a[0], b[0], t) * Eget(Eb[1], a[1], b[1], a[1], b[1], u) *
                               Eget(Eb[2], a[2], b[2], a[2], b[2], v);
        if (eb == 0) continue;
        for (int tau = 0; tau <= lcd[0]; ++tau)
          for (int nu = 0; nu <= lcd[1]; ++nu)
            for (int phi = 0; phi <= lcd[2]; ++phi) {
              const long double ek = Eget(Ek[0], c[0], d[0], c[0], d[0], tau) *
                                     Eget(Ek[1], c[1], d[1], c[1], d[1], nu) * Eget(Ek[2], c[2], d[2], c[2], d[2], phi);
              const long double sgn = ((tau + nu + phi) % 2) ? -1.0L : 1.0L;
              sum += eb * ek * sgn * R[idx(0, t + tau, u + nu, v + phi)];
            }
      }
  return sum;
  };
  // erf(omega r)/r: the Hermite integrals with al_w = al omega^2/(al + omega^2)
  // in place of al, times sqrt(al_w/al) (McMurchie-Davidson with the
  // attenuated kernel); erfc = Coulomb - erf
  long double sum = 0;
  if (g_oper != 1) sum = hsum(al);
  if (g_oper != 0 && g_omega > 0) {
    const long double w2 = (long double)g_omega * g_omega, alw = al * w2 / (al + w2);
    const long double e = sqrtl(alw / al) * hsum(alw);
    sum += g_oper == 1 ? e : -e;
  }
  const long double PI = 3.141592653589793238462643383279502884L;
  return 2 * powl(PI, 2.5L) / (p * q * sqrtl(p + q)) * sum;
}


// ----------------------------------------------------------------------------
// overlap of two contracted shells (for the significant shell-pair list of
// compute_shellpairs, tests/hartree-fock/hartree-fock++.cc:1305-1436), by
// Hermite expansion: S_ij = E^{ij}_0 sqrt(pi/p) per Cartesian direction.
// Returns the Frobenius norm of the (sph-transformed) block.
// ----------------------------------------------------------------------------
static double overlap_block_norm(const Shell& a, const Shell& b) {
  auto ca = cart_comps(a.l), cb = cart_comps(b.l);
  std::vector<double> S(ca.size() * cb.size(), 0.0);
  std::vector<long double> E[3];
  for (int pa = 0; pa < a.nprim(); ++pa)
    for (int pb = 0; pb < b.nprim(); ++pb) {
      const long double al = a.alpha[pa], be = b.alpha[pb], p = al + be;
      for (int d = 0; d < 3; ++d) md_E(a.l, b.l, al, be, (long double)a.O[d] - b.O[d], E[d]);
      const long double pre = powl(3.141592653589793238462643383279502884L / p, 1.5L) * a.coeff[pa] * b.coeff[pb];
      const int T = a.l + b.l + 1;
      for (size_t i = 0; i < ca.size(); ++i)
        for (size_t j = 0; j < cb.size(); ++j) {
          long double v = pre;
          for (int d = 0; d < 3; ++d) v *= E[d][((size_t)ca[i][d] * (b.l + 1) + cb[j][d]) * (T + 1)];
          S[i * cb.size() + j] += (double)v;
        }
    }
  // sph transforms
  std::vector<double> cur = S;
  int n1 = (int)ca.size(), n2 = (int)cb.size();
  if (a.pure && a.l > 0) {
    auto M = sh_matrix(a.l);
    std::vector<double> nx((2 * a.l + 1) * n2, 0.0);
    for (int s = 0; s < 2 * a.l + 1; ++s)
      for (int c = 0; c < n1; ++c)
        for (int j = 0; j < n2; ++j) nx[s * n2 + j] += M[s * n1 + c] * cur[c * n2 + j];
    cur.swap(nx);
    n1 = 2 * a.l + 1;
  }
  if (b.pure && b.l > 0) {
    auto M = sh_matrix(b.l);
    std::vector<double> nx(n1 * (2 * b.l + 1), 0.0);
    for (int i = 0; i < n1; ++i)
      for (int s = 0; s < 2 * b.l + 1; ++s)
        for (int c = 0; c < n2; ++c) nx[i * (2 * b.l + 1) + s] += M[s * n2 + c] * cur[i * n2 + c];
    cur.swap(nx);
  }
  double nrm = 0;
  for (double v : cur) nrm += v * v;
  return std::sqrt(nrm);
}

struct FockSetup {
  std::vector<std::vector<int>> splist;                    // splist[s1] = sorted s2 <= s1
  std::vector<std::vector<std::unique_ptr<ShellPair>>> sp;  // parallel to splist
  std::vector<double> K;                                   // nsh x nsh Schwarz (inf norm)
  std::vector<int> bf1st;
  int nbf = 0;
};

// compute_shellpairs + compute_schwarz_ints (hartree-fock++.cc:1231-1436)
static void fock_setup(const std::vector<Shell>& sh, double threshold, double max_engine_precision, int method,
                       FockSetup& F) {
  const int n = (int)sh.size();
  F.bf1st.resize(n + 1);
  F.bf1st[0] = 0;
  for (int s = 0; s < n; ++s) F.bf1st[s + 1] = F.bf1st[s] + sh[s].size();
  F.nbf = F.bf1st[n];
  F.splist.assign(n, {});
  F.sp.clear();
  F.sp.resize(n);
#pragma omp parallel for schedule(dynamic, 1)
  for (int s1 = 0; s1 < n; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2) {
      const bool same = sh[s1].O == sh[s2].O;
      if (same || overlap_block_norm(sh[s1], sh[s2]) >= threshold) F.splist[s1].push_back(s2);
    }
  // Schwarz matrix, infinity norm, no primitive screening
  F.K.assign((size_t)n * n, 0.0);
#pragma omp parallel
  {
    Engine e;
    e.set_precision(0.0);
#pragma omp for schedule(dynamic, 1)
    for (int s1 = 0; s1 < n; ++s1)
      for (int s2 = 0; s2 <= s1; ++s2) {
        e.compute2(sh[s1], sh[s2], sh[s1], sh[s2]);
        double mx = 0;
        for (double v : e.result) mx = std::max(mx, std::abs(v));
        F.K[(size_t)s1 * n + s2] = F.K[(size_t)s2 * n + s1] = std::sqrt(mx);
      }
  }
  // shell-pair data at max_engine_precision
  const double lnp = std::log(max_engine_precision);
#pragma omp parallel
  {
    Engine e;
    e.set_precision(0.0);
#pragma omp for schedule(dynamic, 1)
    for (int s1 = 0; s1 < n; ++s1) {
      for (int s2 : F.splist[s1]) {
        auto p = std::make_unique<ShellPair>();
        if (method == Original || method == Conservative) {
          p->init(sh[s1], sh[s2], lnp, method);
        } else {
          // primitive Schwarz factors from normalization-free primitives
          // (Shell::extract_primitive(p, false), shell.h:938-947)
          const int n1 = sh[s1].nprim(), n2 = sh[s2].nprim();
          std::vector<double> fac(n1 * n2);
          for (int p1 = 0; p1 < n1; ++p1)
            for (int p2 = 0; p2 < n2; ++p2) {
              Shell a = sh[s1], b = sh[s2];
              a.alpha = {sh[s1].alpha[p1]};
              a.coeff = {1.0};
              a.max_ln_coeff = {0.0};
              b.alpha = {sh[s2].alpha[p2]};
              b.coeff = {1.0};
              b.max_ln_coeff = {0.0};
              double mx = 0, ss = 0;
              if (e.compute2(a, b, a, b))
                for (double v : e.result) {
                  mx = std::max(mx, std::abs(v));
                  ss += v * v;
                }
              fac[p1 * n2 + p2] = std::sqrt(method == SchwarzInf ? mx : std::sqrt(ss));
            }
          p->init_schwarz(sh[s1], sh[s2], lnp, fac.data());
        }
        F.sp[s1].push_back(std::move(p));
      }
    }
  }
}

// compute_2body_fock (hartree-fock++.cc:1574-1772): returns G = 2J - K and
// optionally J and K (J_ab = sum (ab|cd) D_cd, K_ab = sum (ac|bd) D_cd).
// arbitrate: 0 plain (the engine as configured); 1 every quartet is also
// evaluated in long double (11 more significand bits, so the difference
// measures the FP64 evaluation's own error), and a quartet whose two
// evaluations differ by more than 1e-13 (absolute, or relative to its
// largest integral above 1) is evaluated again in quad precision and
// digested from that; 2 every quartet in quad precision.
// narb_out (optional) receives the quartets evaluated in quad precision.
static void fock_build(const std::vector<Shell>& sh, const FockSetup& F, const double* D, double precision,
                       int method, double* G, double* J, double* K, int64_t* nq_out, int arbitrate = 0,
                       int64_t* narb_out = nullptr) {
  const int n = (int)sh.size();
  const int N = F.nbf;
  std::vector<double> Dn((size_t)n * n, 0.0);
  for (int s1 = 0; s1 < n; ++s1)
    for (int s2 = 0; s2 < n; ++s2) {
      double m = 0;
      for (int i = F.bf1st[s1]; i < F.bf1st[s1 + 1]; ++i)
        for (int j = F.bf1st[s2]; j < F.bf1st[s2 + 1]; ++j) m = std::max(m, std::abs(D[(size_t)i * N + j]));
      Dn[(size_t)s1 * n + s2] = m;
    }
  double dmax = 0;
  for (double v : Dn) dmax = std::max(dmax, v);
  const double needed = precision / dmax;
  std::vector<double> jg((size_t)N * N, 0.0), kg((size_t)N * N, 0.0);
  int64_t nq = 0, narb = 0;
#pragma omp parallel reduction(+ : nq, narb)
  {
    Engine e;
    EngineT<long double> e2;
    EngineT<quad> eq;
    e.screening = method;
    e2.screening = method;
    eq.screening = method;
    std::vector<double> vq;
    std::vector<double> jl((size_t)N * N, 0.0), kl((size_t)N * N, 0.0);
#pragma omp for schedule(dynamic, 1)
    for (int s1 = 0; s1 < n; ++s1) {
      if (g_progress) {
#pragma omp critical(progress)
        {
          std::fprintf(stderr, "fock_build: shell row %d of %d started\n", s1 + 1, n);
          std::fflush(stderr);
        }
      }
      const int b1 = F.bf1st[s1], n1 = sh[s1].size();
      for (size_t i2 = 0; i2 < F.splist[s1].size(); ++i2) {
        const int s2 = F.splist[s1][i2];
        const int b2 = F.bf1st[s2], n2 = sh[s2].size();
        const ShellPair* sp12 = F.sp[s1][i2].get();
        const double Dn12 = Dn[(size_t)s1 * n + s2];
        for (int s3 = 0; s3 <= s1; ++s3) {
          const int b3 = F.bf1st[s3], n3 = sh[s3].size();
          const double Dn123 = std::max(std::max(Dn[(size_t)s1 * n + s3], Dn[(size_t)s2 * n + s3]), Dn12);
          const int s4max = (s1 == s3) ? s2 : s3;
          for (size_t i4 = 0; i4 < F.splist[s3].size(); ++i4) {
            const int s4 = F.splist[s3][i4];
            if (s4 > s4max) break;
            if (!g_rows.empty() && !(g_rows[s1] | g_rows[s2] | g_rows[s3] | g_rows[s4])) continue;
            const ShellPair* sp34 = F.sp[s3][i4].get();
            const double Dn1234 =
                std::max(std::max(Dn[(size_t)s1 * n + s4], Dn[(size_t)s2 * n + s4]), std::max(Dn[(size_t)s3 * n + s4], Dn123));
            if (Dn1234 * F.K[(size_t)s1 * n + s2] * F.K[(size_t)s3 * n + s4] < precision) continue;
            const int b4 = F.bf1st[s4], n4 = sh[s4].size();
            const double deg = (s1 == s2 ? 1 : 2) * (s3 == s4 ? 1 : 2) * (s1 == s3 ? (s2 == s4 ? 1 : 2) : 2);
            const double qprec = Dn1234 != 0. ? precision / Dn1234 : needed;
            e.set_precision(qprec);
            if (!e.compute2(sh[s1], sh[s2], sh[s3], sh[s4], sp12, sp34)) continue;
            ++nq;
            const double* buf = e.result.data();
            const size_t nf = (size_t)n1 * n2 * n3 * n4;
            bool redo = arbitrate == 2;
            if (arbitrate == 1) {
              e2.set_precision(qprec);
              e2.compute2(sh[s1], sh[s2], sh[s3], sh[s4], sp12, sp34);
              double dmax = 0, vmax = 1;
              for (size_t f = 0; f < nf; ++f) {
                dmax = std::max(dmax, (double)std::fabs((long double)buf[f] - e2.result[f]));
                vmax = std::max(vmax, std::abs(buf[f]));
              }
              redo = dmax > 1e-13 * vmax;
            }
            if (redo) {
              if (g_progress) {
                static std::atomic<long long> nq_quad{0};
                const long long c = ++nq_quad;
                if (c % 25 == 0) {
#pragma omp critical(progress)
                  {
                    std::fprintf(stderr, "fock_build: %lld quartets in quad precision (row %d)\n", c, s1 + 1);
                    std::fflush(stderr);
                  }
                }
              }
              eq.set_precision(qprec);
              eq.compute2(sh[s1], sh[s2], sh[s3], sh[s4], sp12, sp34);
              vq.resize(nf);
              for (size_t f = 0; f < nf; ++f) vq[f] = (double)eq.result[f];
              buf = vq.data();
              ++narb;
            }
            for (int f1 = 0, f = 0; f1 < n1; ++f1)
              for (int f2 = 0; f2 < n2; ++f2)
                for (int f3 = 0; f3 < n3; ++f3)
                  for (int f4 = 0; f4 < n4; ++f4, ++f) {
                    const size_t i1 = b1 + f1, i2b = b2 + f2, i3 = b3 + f3, i4 = b4 + f4;
                    const double v = buf[f] * deg;
                    jl[i1 * N + i2b] += D[i3 * N + i4] * v;
                    jl[i3 * N + i4] += D[i1 * N + i2b] * v;
                    kl[i1 * N + i3] += D[i2b * N + i4] * v;
                    kl[i2b * N + i4] += D[i1 * N + i3] * v;
                    kl[i1 * N + i4] += D[i2b * N + i3] * v;
                    kl[i2b * N + i3] += D[i1 * N + i4] * v;
                  }
          }
        }
      }
    }
#pragma omp critical
    {
      for (size_t i = 0; i < jg.size(); ++i) {
        jg[i] += jl[i];
        kg[i] += kl[i];
      }
    }
  }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const double Jv = 0.25 * (jg[(size_t)i * N + j] + jg[(size_t)j * N + i]);
      const double Kv = 0.125 * (kg[(size_t)i * N + j] + kg[(size_t)j * N + i]);
      if (J) J[(size_t)i * N + j] = Jv;
      if (K) K[(size_t)i * N + j] = Kv;
      // g = jg - kg / 4, G = (g + g^T) / 2  (hartree-fock++.cc:1730-1766)
      G[(size_t)i * N + j] = 2 * Jv - Kv;
    }
  if (nq_out) *nq_out = nq;
  if (narb_out) *narb_out = narb;
}

}  // namespace orc

// ============================================================================
// C ABI for the Python test harness (ctypes)
// ============================================================================
using namespace orc;

namespace {
struct Basis {
  std::vector<Shell> shells;
};
Basis make_basis(int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                 const double* coeff, const double* origin) {
  Basis b;
  b.shells.resize(nshell);
  size_t off = 0;
  for (int s = 0; s < nshell; ++s) {
    Shell& sh = b.shells[s];
    sh.l = l[s];
    sh.pure = pure[s] != 0;
    sh.alpha.assign(alpha + off, alpha + off + nprim[s]);
    sh.coeff.assign(coeff + off, coeff + off + nprim[s]);
    sh.max_ln_coeff.resize(nprim[s]);
    for (int p = 0; p < nprim[s]; ++p) sh.max_ln_coeff[p] = std::log(std::abs(sh.coeff[p]));
    for (int i = 0; i < 3; ++i) sh.O[i] = origin[3 * s + i];
    off += nprim[s];
  }
  return b;
}
}  // namespace

extern "C" {

int oracle_set_tight_centre(int on) {
  g_tight_centre = on ? 1 : 0;
  return 0;
}

int oracle_set_progress(int on) {
  g_progress = on ? 1 : 0;
  return 0;
}

int oracle_set_operator(int oper, double omega) {
  if (oper < 0 || oper > 2 || omega < 0) return -1;
  g_oper = oper;
  g_omega = omega;
  return 0;
}

int oracle_set_shg_ordering(int o) {
  g_shg_ordering = o;
  return 0;
}

// Boys function, Chebyshev-7 restatement
void oracle_boys_cheb7(int n, const double* T, int mmax, double* out) {
  const auto& b = cheb7();
  for (int i = 0; i < n; ++i) b.eval(out + (size_t)i * (mmax + 1), T[i], mmax);
}
// Boys function, long-double reference (FmEval_Reference2 restatement)
void oracle_boys_reference(int n, const double* T, int mmax, double* out) {
  std::vector<long double> F(mmax + 1);
  for (int i = 0; i < n; ++i) {
    boys_reference((long double)T[i], mmax, F.data());
    for (int m = 0; m <= mmax; ++m) out[(size_t)i * (mmax + 1) + m] = (double)F[m];
  }
}
// Chebyshev table, for cross-checking the library's own table
const double* oracle_boys_table(int* nint, int* mmax) {
  const auto& b = cheb7();
  *nint = Cheb7::NINT;
  *mmax = Cheb7::MMAX;
  return b.c.data();
}

void oracle_shell_renorm(int l, int np, const double* alpha, double* coeff_inout, double* max_ln_coeff,
                         int enforce_unit) {
  renorm(l, np, alpha, coeff_inout, max_ln_coeff, enforce_unit != 0);
}

double oracle_solidharm_coeff(int l, int m, int lx, int ly, int lz) { return sh_coeff(l, m, lx, ly, lz); }

// Engine restatement over a list of quartets (caller shell order). out is
// filled per quartet at offsets[q] (computed here), nonzero[q] = 0 when the
// reference would return nullptr (all primitives screened).  Returns total
// number of doubles written.  OpenMP over quartets, one Engine per thread.
int64_t oracle_eri_quartets(int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                            const double* coeff, const double* origin, int64_t nq, const int32_t* quartets,
                            double precision, int screening, int uniform_cart, double* out, int64_t* offsets,
                            int8_t* nonzero, int nthreads) {
  Basis B = make_basis(nshell, l, pure, nprim, alpha, coeff, origin);
  int64_t total = 0;
  for (int64_t q = 0; q < nq; ++q) {
    offsets[q] = total;
    int64_t sz = 1;
    for (int k = 0; k < 4; ++k) sz *= B.shells[quartets[4 * q + k]].size();
    total += sz;
  }
  if (!out) return total;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    Engine e;
    e.set_precision(precision);
    e.screening = screening;
    e.uniform_cart = uniform_cart != 0;
#pragma omp for schedule(dynamic, 16)
    for (int64_t q = 0; q < nq; ++q) {
      const Shell& s1 = B.shells[quartets[4 * q + 0]];
      const Shell& s2 = B.shells[quartets[4 * q + 1]];
      const Shell& s3 = B.shells[quartets[4 * q + 2]];
      const Shell& s4 = B.shells[quartets[4 * q + 3]];
      const bool nz = e.compute2(s1, s2, s3, s4);
      const int64_t sz = (int64_t)s1.size() * s2.size() * s3.size() * s4.size();
      if (nonzero) nonzero[q] = nz ? 1 : 0;
      if (nz)
        std::memcpy(out + offsets[q], e.result.data(), sizeof(double) * sz);
      else
        std::memset(out + offsets[q], 0, sizeof(double) * sz);
    }
  }
  return total;
}

// The same restatement evaluated in long double throughout (Boys function
// from its series, pair data recomputed from the shells): an arbiter for
// values where the double restatement itself loses digits to cancellation
// (HRR across a distant diffuse-tight pair), far cheaper than the
// McMurchie-Davidson brute force for deep contractions.  Same arguments and
// layout as oracle_eri_quartets; results rounded to double.
int64_t oracle_eri_quartets_ld(int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                               const double* coeff, const double* origin, int64_t nq, const int32_t* quartets,
                               double precision, int screening, int uniform_cart, double* out, int64_t* offsets,
                               int8_t* nonzero, int nthreads) {
  Basis B = make_basis(nshell, l, pure, nprim, alpha, coeff, origin);
  int64_t total = 0;
  for (int64_t q = 0; q < nq; ++q) {
    offsets[q] = total;
    int64_t sz = 1;
    for (int k = 0; k < 4; ++k) sz *= B.shells[quartets[4 * q + k]].size();
    total += sz;
  }
  if (!out) return total;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    EngineT<long double> e;
    e.set_precision(precision);
    e.screening = screening;
    e.uniform_cart = uniform_cart != 0;
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < nq; ++q) {
      const Shell& s1 = B.shells[quartets[4 * q + 0]];
      const Shell& s2 = B.shells[quartets[4 * q + 1]];
      const Shell& s3 = B.shells[quartets[4 * q + 2]];
      const Shell& s4 = B.shells[quartets[4 * q + 3]];
      const bool nz = e.compute2(s1, s2, s3, s4);
      const int64_t sz = (int64_t)s1.size() * s2.size() * s3.size() * s4.size();
      if (nonzero) nonzero[q] = nz ? 1 : 0;
      for (int64_t k = 0; k < sz; ++k) out[offsets[q] + k] = nz ? (double)e.result[k] : 0.0;
    }
  }
  return total;
}

// The same in quad precision (__float128).
int64_t oracle_eri_quartets_q(int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                               const double* coeff, const double* origin, int64_t nq, const int32_t* quartets,
                               double precision, int screening, int uniform_cart, double* out, int64_t* offsets,
                               int8_t* nonzero, int nthreads) {
  Basis B = make_basis(nshell, l, pure, nprim, alpha, coeff, origin);
  int64_t total = 0;
  for (int64_t q = 0; q < nq; ++q) {
    offsets[q] = total;
    int64_t sz = 1;
    for (int k = 0; k < 4; ++k) sz *= B.shells[quartets[4 * q + k]].size();
    total += sz;
  }
  if (!out) return total;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    EngineT<quad> e;
    e.set_precision(precision);
    e.screening = screening;
    e.uniform_cart = uniform_cart != 0;
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < nq; ++q) {
      const Shell& s1 = B.shells[quartets[4 * q + 0]];
      const Shell& s2 = B.shells[quartets[4 * q + 1]];
      const Shell& s3 = B.shells[quartets[4 * q + 2]];
      const Shell& s4 = B.shells[quartets[4 * q + 3]];
      const bool nz = e.compute2(s1, s2, s3, s4);
      const int64_t sz = (int64_t)s1.size() * s2.size() * s3.size() * s4.size();
      if (nonzero) nonzero[q] = nz ? 1 : 0;
      for (int64_t k = 0; k < sz; ++k) out[offsets[q] + k] = nz ? (double)e.result[k] : 0.0;
    }
  }
  return total;
}

// Brute-force McMurchie-Davidson (long double) for a list of quartets in
// caller order; no screening.
int64_t oracle_eri_brute(int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                         const double* coeff, const double* origin, int64_t nq, const int32_t* quartets,
                         double* out, int nthreads) {
  Basis B = make_basis(nshell, l, pure, nprim, alpha, coeff, origin);
  std::vector<int64_t> off(nq + 1, 0);
  for (int64_t q = 0; q < nq; ++q) {
    int64_t sz = 1;
    for (int k = 0; k < 4; ++k) sz *= B.shells[quartets[4 * q + k]].size();
    off[q + 1] = off[q] + sz;
  }
  if (!out) return off[nq];
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < nq; ++q) {
    const Shell* s[4];
    for (int k = 0; k < 4; ++k) s[k] = &B.shells[quartets[4 * q + k]];
    int nc[4];
    std::vector<std::array<int, 3>> cc[4];
    for (int k = 0; k < 4; ++k) {
      cc[k] = cart_comps(s[k]->l);
      nc[k] = (int)cc[k].size();
    }
    long double O[4][3];
    for (int k = 0; k < 4; ++k)
      for (int i = 0; i < 3; ++i) O[k][i] = s[k]->O[i];
    std::vector<long double> cart((size_t)nc[0] * nc[1] * nc[2] * nc[3], 0.0L);
    for (int p0 = 0; p0 < s[0]->nprim(); ++p0)
      for (int p1 = 0; p1 < s[1]->nprim(); ++p1)
        for (int p2 = 0; p2 < s[2]->nprim(); ++p2)
          for (int p3 = 0; p3 < s[3]->nprim(); ++p3) {
            const long double cf = (long double)s[0]->coeff[p0] * s[1]->coeff[p1] * s[2]->coeff[p2] * s[3]->coeff[p3];
            size_t idx = 0;
            for (int i0 = 0; i0 < nc[0]; ++i0)
              for (int i1 = 0; i1 < nc[1]; ++i1)
                for (int i2 = 0; i2 < nc[2]; ++i2)
                  for (int i3 = 0; i3 < nc[3]; ++i3, ++idx)
                    cart[idx] += cf * md_primitive(cc[0][i0].data(), cc[1][i1].data(), cc[2][i2].data(),
                                                   cc[3][i3].data(), s[0]->alpha[p0], s[1]->alpha[p1],
                                                   s[2]->alpha[p2], s[3]->alpha[p3], O[0], O[1], O[2], O[3]);
          }
    // sph transforms in caller order
    std::vector<long double> cur = cart;
    int dims[4] = {nc[0], nc[1], nc[2], nc[3]};
    for (int k = 0; k < 4; ++k) {
      if (!s[k]->pure) continue;
      const int lk = s[k]->l;
      auto M = sh_matrix(lk);
      const int ncc = ncart(lk), ns = 2 * lk + 1;
      int pre = 1, post = 1;
      for (int j = 0; j < k; ++j) pre *= dims[j];
      for (int j = k + 1; j < 4; ++j) post *= dims[j];
      std::vector<long double> nx((size_t)pre * ns * post, 0.0L);
      for (int a = 0; a < pre; ++a)
        for (int sidx = 0; sidx < ns; ++sidx)
          for (int c = 0; c < ncc; ++c) {
            const long double v = M[sidx * ncc + c];
            if (v == 0.0L) continue;
            for (int r = 0; r < post; ++r)
              nx[((size_t)a * ns + sidx) * post + r] += v * cur[((size_t)a * ncc + c) * post + r];
          }
      cur.swap(nx);
      dims[k] = ns;
    }
    for (size_t i = 0; i < cur.size(); ++i) out[off[q] + i] = (double)cur[i];
  }
  return off[nq];
}

// Screened Fock build restating compute_shellpairs + compute_schwarz_ints +
// compute_2body_fock.  D, G, J, K are nbf x nbf row-major; J/K may be NULL.
// splist_out (optional, nsh*nsh bytes) receives the significant-pair mask.
int oracle_fock_2e(int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                   const double* coeff, const double* origin, const double* D, double precision, double* G,
                   int method, double* JK /* 2*nbf*nbf or NULL */, int nthreads, int64_t* nq) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  Basis B = make_basis(nshell, l, pure, nprim, alpha, coeff, origin);
  FockSetup F;
  fock_setup(B.shells, 1e-12, std::numeric_limits<double>::epsilon() / 1e10, method, F);
  const size_t NN = (size_t)F.nbf * F.nbf;
  fock_build(B.shells, F, D, precision, method, G, JK ? JK : nullptr, JK ? JK + NN : nullptr, nq);
  return F.nbf;
}

// oracle_fock_2e with extended-precision arbitration of the ill-conditioned
// quartets (fock_build's arbitrate 1) or every quartet in quad precision
// (arbitrate 2); narb receives the quartets evaluated in quad precision.
int oracle_fock_2e_arbitrated(int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                              const double* coeff, const double* origin, const double* D, double precision,
                              double* G, int method, double* JK, int nthreads, int arbitrate, int64_t* nq,
                              int64_t* narb) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  Basis B = make_basis(nshell, l, pure, nprim, alpha, coeff, origin);
  FockSetup F;
  fock_setup(B.shells, 1e-12, std::numeric_limits<double>::epsilon() / 1e10, method, F);
  const size_t NN = (size_t)F.nbf * F.nbf;
  fock_build(B.shells, F, D, precision, method, G, JK ? JK : nullptr, JK ? JK + NN : nullptr, nq, arbitrate, narb);
  return F.nbf;
}

// Restricts the Fock builds to quartets touching the shells with mask[s]
// != 0 (mask NULL or nshell 0: every quartet).
void oracle_set_fock_rows(int nshell, const int8_t* mask) {
  if (!mask || nshell <= 0)
    g_rows.clear();
  else
    g_rows.assign(mask, mask + nshell);
}

// Significant shell pairs (overlap-norm >= threshold or same centre) and the
// Schwarz matrix, as the Fock set-up computes them.
int oracle_fock_pairs(int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                      const double* coeff, const double* origin, double threshold, int8_t* mask, double* K) {
  Basis B = make_basis(nshell, l, pure, nprim, alpha, coeff, origin);
  FockSetup F;
  fock_setup(B.shells, threshold, std::numeric_limits<double>::epsilon() / 1e10, Original, F);
  std::memset(mask, 0, (size_t)nshell * nshell);
  for (int s1 = 0; s1 < nshell; ++s1)
    for (int s2 : F.splist[s1]) mask[(size_t)s1 * nshell + s2] = 1;
  std::memcpy(K, F.K.data(), sizeof(double) * nshell * nshell);
  return 0;
}

}  // extern "C"
