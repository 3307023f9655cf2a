// Host side of the C ABI declared in include/l2g.h: device context, Boys
// table, basis upload, shell-pair data, batch planning and kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <deque>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/l2g.h"
#include "eri_device.h"

extern const L2gClassInfo l2g_gen_classes[];
extern const int l2g_gen_nclasses;
extern const L2gClassInfo l2g_gen_schwarz[];
extern const int l2g_gen_nschwarz;
extern const L2gClassInfo l2g_gen_fock[];
extern const int l2g_gen_nfock;
extern const L2gClassInfo l2g_gen_classes32[];
extern const int l2g_gen_nclasses32;
extern const L2gClassInfo l2g_gen_fock32[];
extern const int l2g_gen_nfock32;
extern const int l2g_gen_lmax;

namespace {

thread_local std::string g_err;

struct L2gError : std::runtime_error {
  int code;
  L2gError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess)                                                                           \
      throw L2gError(e_ == hipErrorOutOfMemory ? L2G_E_OOM : L2G_E_NODEVICE,                      \
                     std::string(#x) + ": " + hipGetErrorString(e_));                              \
  } while (0)

template <class F>
int guarded(F&& f) {
  try {
    f();
    return L2G_OK;
  } catch (const L2gError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return L2G_E_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return L2G_E_INVALID;
  }
}

int ncart(int l) { return (l + 1) * (l + 2) / 2; }
int nfun(int l, int pure) { return pure ? 2 * l + 1 : ncart(l); }

// ---------------------------------------------------------------------------
// Boys table: for every interval [iv/7, (iv+1)/7] and m, the degree-7
// polynomial in xd = 7T - iv - 1/2 that interpolates F_m(T) at the 8
// Chebyshev points of the interval.  F_m at the points comes from the
// Maclaurin series of F_40 (all terms positive, long double) and downward
// recursion.  Same construction as FmEval_Chebyshev7's table
// (include/libint2/boys.h:252-482, parameters in
// include/libint2/util/boys_cheb.nb: delta = 1/7, order 7, m <= 40).
// ---------------------------------------------------------------------------
struct BoysTable {
  std::vector<double> t;  // [m][iv][8]
  BoysTable() {
    const int NI = L2G_BOYS_NINT, MM = L2G_BOYS_MMAX, NC = 8;
    t.assign((size_t)(MM + 1) * NI * NC, 0.0);
    const long double pi = 3.14159265358979323846264338327950288L;
    long double node[NC];
    for (int k = 0; k < NC; ++k) node[k] = cosl(pi * (2 * k + 1) / (2.0L * NC));  // in [-1, 1]
    // Chebyshev polynomial monomial coefficients, T_j(u) = sum_i P[j][i] u^i
    long double P[NC][NC];
    std::memset(P, 0, sizeof(P));
    P[0][0] = 1;
    P[1][1] = 1;
    for (int j = 2; j < NC; ++j)
      for (int i = 0; i < NC; ++i) P[j][i] = (i ? 2 * P[j - 1][i - 1] : 0.0L) - P[j - 2][i];
    long double vals[MM + 1][NC];
    for (int iv = 0; iv < NI; ++iv) {
      for (int k = 0; k < NC; ++k) {
        const long double T = ((long double)iv + 0.5L + 0.5L * node[k]) / 7.0L;
        // F_MM(T) = e^{-T} sum_i T^i / prod_{j=0..i} (MM + j + 1/2) / 2
        long double term = 1.0L / (MM + 0.5L), sum = term;
        for (int i = 1; i < 2000; ++i) {
          term *= T / (MM + i + 0.5L);
          sum += term;
          if (term < sum * 1e-21L) break;
        }
        const long double eT = expl(-T);
        long double f = 0.5L * eT * sum;
        vals[MM][k] = f;
        for (int m = MM; m > 0; --m) {
          f = (2 * T * f + eT) / (2 * m - 1);
          vals[m - 1][k] = f;
        }
      }
      for (int m = 0; m <= MM; ++m) {
        long double c[NC];
        for (int j = 0; j < NC; ++j) {
          long double s = 0;
          for (int k = 0; k < NC; ++k) s += vals[m][k] * cosl(pi * j * (2 * k + 1) / (2.0L * NC));
          c[j] = s * (j ? 2.0L : 1.0L) / NC;
        }
        double* out = &t[((size_t)m * NI + iv) * NC];
        long double scale = 1;  // u = 2 xd
        for (int i = 0; i < NC; ++i) {
          long double s = 0;
          for (int j = 0; j < NC; ++j) s += c[j] * P[j][i];
          out[i] = (double)(s * scale);
          scale *= 2;
        }
      }
    }
  }
};
const BoysTable& boys_table() {
  static BoysTable tab;
  return tab;
}

// ---------------------------------------------------------------------------
// normalization and solid harmonics
// ---------------------------------------------------------------------------
double double_factorial_km1(int k) {  // (k-1)!!
  double r = 1;
  for (int i = k - 1; i > 1; i -= 2) r *= i;
  return r;
}
double factorial(int n) {
  double r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
double choose(int n, int k) {
  if (k < 0 || k > n) return 0;
  return factorial(n) / (factorial(k) * factorial(n - k));
}
int sgn_parity(int i) { return (i & 1) ? -1 : 1; }

// ---------------------------------------------------------------------------
// objects behind the opaque handles
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t b) {
    if (p && bytes >= b) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (b == 0) return;
    HIPCHK(hipMalloc(&p, b));
    bytes = b;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

template <class T>
void upload(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  d.alloc(std::max<size_t>(h.size() * sizeof(T), 16));
  if (!h.empty()) HIPCHK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

}  // namespace

struct l2g_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  DevBuf boys;
  int shg_gaussian = 0;
  int ncu = 256;
  int class_of[8][8][8][8];  // index into l2g_gen_classes
  int fock_of[8][8][8][8];   // index into l2g_gen_fock (its own order)
  int class32_of[8][8][8][8];  // index into l2g_gen_classes32
  int fock32_of[8][8][8][8];   // index into l2g_gen_fock32
  int schwarz_of[8][8];
  // bases, plans and Fock builders created on this context and not yet
  // destroyed; l2g_finalize refuses while any is alive
  std::atomic<int> live{0};
};

struct l2g_basis_s {
  l2g_context_s* ctx;
  int nshell;
  std::vector<int> l, pure, nprim, off;
  std::vector<double> alpha, coeff, lnc, origin;
  std::unordered_map<uint64_t, bool> orient;  // second_is_centre() memo

  ~l2g_basis_s() {
    if (ctx) --ctx->live;
  }

  // Which shell of a pair with equal l is the HRR build centre.  The
  // reference keeps the caller's order for equal l (engine.impl.h:1181-1208);
  // the integral does not depend on the choice, its rounding does: HRR
  // (a,b) = sum_k C(b,k) AB^k (a+k,0) cancels badly when the build centre is
  // the diffuse shell of a distant pair, because the pair's charge sits near
  // the tight shell.  True when shell j is tighter, i.e. when the
  // K-weighted mean of |P - B_j| = alpha_i/gamma |AB| is below that of
  // |P - A_i| = alpha_j/gamma |AB|.
  bool second_is_centre(int i, int j) {
    if (i == j || l[i] != l[j]) return false;
    const uint64_t key = (uint64_t)(uint32_t)i << 32 | (uint32_t)j;
    auto it = orient.find(key);
    if (it != orient.end()) return it->second;
    double AB2 = 0;
    for (int k = 0; k < 3; ++k) AB2 += (origin[3 * i + k] - origin[3 * j + k]) * (origin[3 * i + k] - origin[3 * j + k]);
    double s = 0;
    if (AB2 > 0)
      for (int p1 = 0; p1 < nprim[i]; ++p1)
        for (int p2 = 0; p2 < nprim[j]; ++p2) {
          const double a1 = alpha[off[i] + p1], a2 = alpha[off[j] + p2], g = a1 + a2;
          const double w = std::fabs(coeff[off[i] + p1] * coeff[off[j] + p2]) * std::exp(-a1 * a2 / g * AB2) / g;
          s += w * (a2 - a1) / g;
        }
    const bool r = s > 0;
    orient.emplace(key, r);
    orient.emplace((uint64_t)(uint32_t)j << 32 | (uint32_t)i, s < 0);
    return r;
  }
  // canonical bra orientation of (i, j): l_i >= l_j, ties by second_is_centre
  bool swap_pair(int i, int j) { return l[i] < l[j] || (l[i] == l[j] && second_is_centre(i, j)); }
};

struct ClassWork {
  int cls;  // index into l2g_gen_classes (l2g_gen_classes32 when fp32)
  bool fp32 = false;
  int64_t first, count;
  int64_t nprimq;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float ms = 0;
};

struct l2g_eri_plan_s {
  l2g_basis_s* basis;
  int64_t nq = 0;
  int64_t total = 0;
  std::vector<int64_t> offsets;
  std::vector<ClassWork> classes;
  double precision, ln_precision, scale;
  int method, uniform;
  double fp32_threshold = 0;  // quartets with K12 K34 below it run in FP32
  // device arrays
  DevBuf bra, ket, off, meta, qidx;
  DevBuf pair_first, pair_np, pair_AB, ppP, ppPA, ppAux, ppNs;
  DevBuf out, ybuf, nz;
  int64_t ystride = 0;
  int grid = 0, block = 128;
  ~l2g_eri_plan_s() {
    for (auto& c : classes) {
      if (c.ev0) (void)hipEventDestroy(c.ev0);
      if (c.ev1) (void)hipEventDestroy(c.ev1);
    }
  }
};

// FP32 selection of a mixed batch: a quartet runs in FP32 when its Schwarz
// bound K_ab K_cd is below the caller's threshold and its pre-HRR bound
// S_ab S_cd (schwarz_impl) is below this factor times the threshold, which
// keeps eps_fp32 S_ab S_cd, the scale of the FP32 rounding that reaches an
// integral, near eps_fp32 times the threshold.
#ifndef L2G_FP32_HRR_FACTOR
#define L2G_FP32_HRR_FACTOR 16.0
#endif

namespace {

void schwarz_impl(l2g_basis B, int64_t npairs, const int32_t* pairs, double precision, int norm, double* K,
                  double* S);

}  // namespace

namespace {

// Shell-pair data for ordered pair (i, j), restating ShellPair::init
// (include/libint2/shell.h:1137-1256) for Original and Conservative
// screening, with the pair's contraction coefficients folded into K.
struct PairBuilder {
  const l2g_basis_s& B;
  double ln_prec;
  int method;
  std::vector<int> first, np;
  std::vector<double> AB;
  std::vector<double4> P, PA;
  std::vector<double2> aux;
  std::vector<double> ns;
  std::unordered_map<uint64_t, int> id;

  // Schwarz / SchwarzInf primitive screening (shell.h:1258-1328): factors of
  // the ordered pair (i, j), p1-major, supplied per pair by the caller
  std::unordered_map<uint64_t, const std::vector<double>*> pfac;

  PairBuilder(const l2g_basis_s& b, double lp, int m) : B(b), ln_prec(lp), method(m) {}

  int get(int i, int j) {
    const uint64_t key = (uint64_t)(uint32_t)i << 32 | (uint32_t)j;
    auto it = id.find(key);
    if (it != id.end()) return it->second;
    const std::vector<double>* fac = nullptr;
    bool fac_t = false;
    if (method == L2G_SCREEN_SCHWARZ || method == L2G_SCREEN_SCHWARZINF) {
      auto f = pfac.find(key);
      if (f == pfac.end()) {
        f = pfac.find((uint64_t)(uint32_t)j << 32 | (uint32_t)i);
        fac_t = true;
      }
      if (f == pfac.end()) throw L2gError(L2G_E_STATE, "no primitive Schwarz factors for a shell pair");
      fac = f->second;
    }
    const int pid = (int)first.size();
    id.emplace(key, pid);
    first.push_back((int)P.size());
    const double* A = &B.origin[3 * i];
    const double* Bo = &B.origin[3 * j];
    double ab[3], AB2 = 0;
    for (int k = 0; k < 3; ++k) {
      ab[k] = A[k] - Bo[k];
      AB2 += ab[k] * ab[k];
      AB.push_back(ab[k]);
    }
    const int n1 = B.nprim[i], n2 = B.nprim[j];
    const int l1 = B.l[i], l2 = B.l[j];
    int cnt = 0;
    for (int p1 = 0; p1 < n1; ++p1)
      for (int p2 = 0; p2 < n2; ++p2) {
        const double a1 = B.alpha[B.off[i] + p1], a2 = B.alpha[B.off[j] + p2];
        const double gamma = a1 + a2, oog = 1.0 / gamma;
        const double rho = a1 * a2 * oog;
        const double mr = -rho * AB2;
        double ln_scr = mr + B.lnc[B.off[i] + p1] + B.lnc[B.off[j] + p2];
        if (fac) {
          const double f = fac_t ? (*fac)[(size_t)p2 * n1 + p1] : (*fac)[(size_t)p1 * n2 + p2];
          ln_scr = std::log((double)(n1 * n2) * f) + B.lnc[B.off[i] + p1] + B.lnc[B.off[j] + p2];
          if (ln_scr < ln_prec) continue;
        }
        if (method == L2G_SCREEN_ORIGINAL && ln_scr < ln_prec) continue;
        double Pp[3];
        for (int k = 0; k < 3; ++k) Pp[k] = (AB2 == 0.0) ? A[k] : (a1 * A[k] + a2 * Bo[k]) * oog;
        double nonsph = 0;
        if (method == L2G_SCREEN_CONSERVATIVE) {
          double mpa = 0, mpb = 0;
          for (int k = 0; k < 3; ++k) {
            mpa = std::max(mpa, std::fabs(Pp[k] - A[k]));
            mpb = std::max(mpb, std::fabs(Pp[k] - Bo[k]));
          }
          nonsph = std::max(std::pow(mpa, l1) * std::pow(mpb, l2),
                            factorial(l1) * factorial(l2) * std::pow(oog, l1 + l2));
          ln_scr += 1.777485947591722872387900 + std::log(oog) + std::log(std::max(nonsph, 1.0)) +
                    std::log((double)(n1 * n2));
          if (ln_scr < ln_prec) continue;
        }
        const double K = 5.9149671727956128778 * std::exp(mr) * oog;
        const double c = B.coeff[B.off[i] + p1] * B.coeff[B.off[j] + p2];
        P.push_back(double4{Pp[0], Pp[1], Pp[2], gamma});
        PA.push_back(double4{Pp[0] - A[0], Pp[1] - A[1], Pp[2] - A[2], K * c});
        aux.push_back(double2{ln_scr, 0.5 * oog});
        ns.push_back(nonsph);
        ++cnt;
      }
    np.push_back(cnt);
    return pid;
  }
};

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* l2g_version(void) { return "libint_amd 0.1 (gfx950)"; }

const char* l2g_last_error(void) { return g_err.c_str(); }

int l2g_max_l(void) { return l2g_gen_lmax; }

int l2g_shell_renorm(int l, int nprim, const double* alpha, double* coeff, double* max_ln_coeff, int enforce_unit) {
  return guarded([&] {
    if (l < 0 || nprim <= 0 || !alpha || !coeff) throw std::invalid_argument("l2g_shell_renorm: bad arguments");
    const double pi32 = 5.56832799683170784528481798212;  // pi^(3/2)
    const double dfl = double_factorial_km1(2 * l);
    for (int p = 0; p < nprim; ++p) {
      if (alpha[p] < 0) throw std::invalid_argument("l2g_shell_renorm: negative exponent");
      if (alpha[p] != 0) {
        const double ta = 2 * alpha[p];
        coeff[p] *= std::sqrt(std::pow(2.0, l) * std::pow(ta, l + 1) * std::sqrt(ta) / (pi32 * dfl));
      }
    }
    if (enforce_unit) {
      double s = 0;
      for (int p = 0; p < nprim; ++p)
        for (int q = 0; q <= p; ++q) {
          const double g = alpha[p] + alpha[q];
          s += (p == q ? 1 : 2) * dfl * pi32 * coeff[p] * coeff[q] / (std::pow(2.0, l) * std::pow(g, l + 1) * std::sqrt(g));
        }
      const double f = 1.0 / std::sqrt(s);
      for (int p = 0; p < nprim; ++p) coeff[p] *= f;
    }
    if (max_ln_coeff)
      for (int p = 0; p < nprim; ++p) max_ln_coeff[p] = std::log(std::fabs(coeff[p]));
  });
}

double l2g_solid_harmonics_coeff(int l, int m, int lx, int ly, int lz) {
  const int am = std::abs(m);
  if ((lx + ly - am) % 2) return 0.0;
  const int j = (lx + ly - am) / 2;
  if (j < 0) return 0.0;
  const int comp = (m >= 0) ? 1 : -1;
  const int i = am - lx;
  if (comp != sgn_parity(std::abs(i))) return 0.0;
  double pfac = std::sqrt(((factorial(2 * lx) * factorial(2 * ly) * factorial(2 * lz)) / factorial(2 * l)) *
                          (factorial(l - am) / factorial(l)) * (1.0 / factorial(l + am)) *
                          (1.0 / (factorial(lx) * factorial(ly) * factorial(lz))));
  pfac /= (double)(1L << l);
  pfac *= (m < 0) ? sgn_parity((i - 1) / 2) : sgn_parity(i / 2);
  double sum = 0;
  for (int ii = j; ii <= (l - am) / 2; ++ii) {
    double t = choose(l, ii) * choose(ii, j) * (sgn_parity(ii) * factorial(2 * (l - ii))) / factorial(l - am - 2 * ii);
    double s1 = 0;
    for (int k = std::max((lx - am) / 2, 0); k <= std::min(j, lx / 2); ++k)
      if (lx - 2 * k <= am) s1 += choose(j, k) * choose(am, lx - 2 * k) * sgn_parity(k);
    sum += t * s1;
  }
  sum *= std::sqrt(double_factorial_km1(2 * l) /
                   (double_factorial_km1(2 * lx) * double_factorial_km1(2 * ly) * double_factorial_km1(2 * lz)));
  return (m == 0) ? pfac * sum : std::sqrt(2.0) * pfac * sum;
}

int l2g_boys_table(double* out, int* nint, int* mmax) {
  return guarded([&] {
    if (nint) *nint = L2G_BOYS_NINT;
    if (mmax) *mmax = L2G_BOYS_MMAX;
    if (out) {
      const auto& t = boys_table().t;
      std::memcpy(out, t.data(), t.size() * sizeof(double));
    }
  });
}

int l2g_init(int device, l2g_context* out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("l2g_init: null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw L2gError(L2G_E_NODEVICE, "no HIP device available");
    if (device < 0 || device >= n) throw std::invalid_argument("l2g_init: device index out of range");
    auto ctx = std::make_unique<l2g_context_s>();
    ctx->device = device;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    ctx->ncu = prop.multiProcessorCount;
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    const auto& t = boys_table().t;
    ctx->boys.alloc(t.size() * sizeof(double));
    HIPCHK(hipMemcpy(ctx->boys.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    for (auto& a : ctx->class_of)
      for (auto& b : a)
        for (auto& c : b)
          for (auto& d : c) d = -1;
    for (int i = 0; i < l2g_gen_nclasses; ++i) {
      const auto& c = l2g_gen_classes[i];
      ctx->class_of[c.la][c.lb][c.lc][c.ld] = i;
    }
    auto fill = [](int (&m)[8][8][8][8], const L2gClassInfo* tab, int n) {
      for (auto& a : m)
        for (auto& b : a)
          for (auto& c : b)
            for (auto& d : c) d = -1;
      for (int i = 0; i < n; ++i)
        if (tab[i].launch) m[tab[i].la][tab[i].lb][tab[i].lc][tab[i].ld] = i;
    };
    fill(ctx->fock_of, l2g_gen_fock, l2g_gen_nfock);
    fill(ctx->class32_of, l2g_gen_classes32, l2g_gen_nclasses32);
    fill(ctx->fock32_of, l2g_gen_fock32, l2g_gen_nfock32);
    for (auto& a : ctx->schwarz_of)
      for (auto& b : a) b = -1;
    for (int i = 0; i < l2g_gen_nschwarz; ++i) {
      const auto& c = l2g_gen_schwarz[i];
      if (c.launch) ctx->schwarz_of[c.la][c.lb] = i;
    }
    *out = ctx.release();
  });
}

int l2g_finalize(l2g_context ctx) {
  return guarded([&] {
    if (!ctx) return;
    if (ctx->live > 0)
      throw L2gError(L2G_E_STATE, std::to_string(ctx->live.load()) +
                                      " basis/plan/Fock object(s) of this context are still alive; destroy them first");
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int l2g_set_solid_harmonics_ordering(l2g_context ctx, int ordering) {
  return guarded([&] {
    if (!ctx) throw L2gError(L2G_E_STATE, "null context");
    if (ordering != L2G_SHG_STANDARD && ordering != L2G_SHG_GAUSSIAN)
      throw std::invalid_argument("solid harmonics ordering value not recognized");
    ctx->shg_gaussian = (ordering == L2G_SHG_GAUSSIAN);
  });
}

int l2g_basis_create(l2g_context ctx, int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                     const double* coeff, const double* origin, l2g_basis* out) {
  return guarded([&] {
    if (!ctx) throw L2gError(L2G_E_STATE, "null context");
    if (!out || nshell < 0 || (nshell > 0 && (!l || !pure || !nprim || !alpha || !coeff || !origin)))
      throw std::invalid_argument("l2g_basis_create: bad arguments");
    auto b = std::make_unique<l2g_basis_s>();
    b->ctx = ctx;
    ++ctx->live;
    b->nshell = nshell;
    b->l.assign(l, l + nshell);
    b->pure.assign(pure, pure + nshell);
    b->nprim.assign(nprim, nprim + nshell);
    b->off.resize(nshell + 1);
    b->off[0] = 0;
    for (int s = 0; s < nshell; ++s) {
      if (l[s] < 0) throw std::invalid_argument("negative angular momentum");
      if (l[s] > l2g_gen_lmax)
        throw L2gError(L2G_E_LMAX, "angular momentum " + std::to_string(l[s]) + " exceeds the compiled limit " +
                                       std::to_string(l2g_gen_lmax));
      if (nprim[s] <= 0) throw std::invalid_argument("shell without primitives");
      b->off[s + 1] = b->off[s] + nprim[s];
    }
    const int np = b->off[nshell];
    b->alpha.assign(alpha, alpha + np);
    b->coeff.assign(coeff, coeff + np);
    b->lnc.resize(np);
    for (int p = 0; p < np; ++p) b->lnc[p] = std::log(std::fabs(b->coeff[p]));
    b->origin.assign(origin, origin + 3 * nshell);
    *out = b.release();
  });
}

int l2g_basis_destroy(l2g_basis b) {
  delete b;
  return L2G_OK;
}

static int plan_create_impl(l2g_basis B, int64_t nq, const int32_t* quartets, double precision, int screening,
                            double scale, int uniform_cart, double fp32_threshold, l2g_eri_plan* out) {
  return guarded([&] {
    if (!B) throw L2gError(L2G_E_STATE, "null basis");
    if (!out || nq < 0 || (nq > 0 && !quartets)) throw std::invalid_argument("l2g_eri_plan_create: bad arguments");
    if (screening != L2G_SCREEN_ORIGINAL && screening != L2G_SCREEN_CONSERVATIVE)
      throw std::invalid_argument(
          "without precomputed shell pair data only the Original or Conservative screening methods are available");
    l2g_context_s* ctx = B->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    auto plan = std::make_unique<l2g_eri_plan_s>();
    plan->basis = B;
    plan->nq = nq;
    plan->scale = scale;
    plan->method = screening;
    plan->uniform = uniform_cart ? 1 : 0;
    if (precision <= 0) {
      plan->precision = 0;
      plan->ln_precision = std::numeric_limits<double>::lowest();
    } else {
      plan->precision = precision;
      plan->ln_precision = std::log(precision);
    }
    plan->fp32_threshold = fp32_threshold;
    PairBuilder pb(*B, plan->ln_precision, screening);
    // Schwarz and FP32-selection factors of the canonical pairs, when the
    // batch is mixed
    std::unordered_map<uint64_t, std::pair<double, double>> Kp;
    auto pkey = [&](int a, int b) {
      if (B->swap_pair(a, b)) std::swap(a, b);
      return (uint64_t)(uint32_t)a << 32 | (uint32_t)b;
    };
    if (fp32_threshold > 0) {
      std::vector<int32_t> pr;
      for (int64_t q = 0; q < nq; ++q)
        for (int h = 0; h < 2; ++h) {
          const int a = quartets[4 * q + 2 * h], b = quartets[4 * q + 2 * h + 1];
          if (a < 0 || b < 0 || a >= B->nshell || b >= B->nshell) throw std::invalid_argument("shell index out of range");
          if (Kp.emplace(pkey(a, b), std::make_pair(0.0, 0.0)).second) {
            pr.push_back(a);
            pr.push_back(b);
          }
        }
      std::vector<double> K(pr.size() / 2), S(pr.size() / 2);
      schwarz_impl(B, (int64_t)K.size(), pr.data(), 0.0, 0, K.data(), S.data());
      for (size_t i = 0; i < K.size(); ++i) Kp[pkey(pr[2 * i], pr[2 * i + 1])] = {K[i], S[i]};
    }
    // FP32 when both the Schwarz bound and the pre-HRR bound are small
    auto use_fp32 = [&](const int32_t* s) {
      const auto& x = Kp[pkey(s[0], s[1])];
      const auto& y = Kp[pkey(s[2], s[3])];
      return x.first * y.first < fp32_threshold && x.second * y.second < L2G_FP32_HRR_FACTOR * fp32_threshold;
    };
    const int nc64 = l2g_gen_nclasses;
    // canonicalise every quartet and bucket by class
    struct Item {
      int bp, kp, meta;
      int64_t off, q;
      int64_t w;
    };
    std::vector<std::vector<Item>> buckets(l2g_gen_nclasses + l2g_gen_nclasses32);
    plan->offsets.resize(nq);
    int64_t total = 0;
    for (int64_t q = 0; q < nq; ++q) {
      const int32_t* s = quartets + 4 * q;
      for (int k = 0; k < 4; ++k)
        if (s[k] < 0 || s[k] >= B->nshell) throw std::invalid_argument("shell index out of range");
      const int l1 = B->l[s[0]], l2 = B->l[s[1]], l3 = B->l[s[2]], l4 = B->l[s[3]];
      // engine.impl.h:1181-1208, equal-l pairs oriented by second_is_centre
      const bool stb = B->swap_pair(s[0], s[1]), stk = B->swap_pair(s[2], s[3]), sbk = (l1 + l2) > (l3 + l4);
      const int b1 = sbk ? (stk ? s[3] : s[2]) : (stb ? s[1] : s[0]);
      const int b2 = sbk ? (stk ? s[2] : s[3]) : (stb ? s[0] : s[1]);
      const int k1 = sbk ? (stb ? s[1] : s[0]) : (stk ? s[3] : s[2]);
      const int k2 = sbk ? (stb ? s[0] : s[1]) : (stk ? s[2] : s[3]);
      int cls = ctx->class_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
      if (cls < 0) throw L2gError(L2G_E_LMAX, "no kernel for this angular momentum class");
      if (fp32_threshold > 0 && use_fp32(s)) {
        const int c32 = ctx->class32_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
        if (c32 >= 0) cls = nc64 + c32;
      }
      auto purebit = [&](int sh) { return (B->pure[sh] && B->l[sh] > 0) ? 1 : 0; };
      const int meta = (stb ? 1 : 0) | (stk ? 2 : 0) | (sbk ? 4 : 0) | (purebit(b1) << 3) | (purebit(b2) << 4) |
                       (purebit(k1) << 5) | (purebit(k2) << 6);
      Item it;
      it.bp = pb.get(b1, b2);
      it.kp = pb.get(k1, k2);
      it.meta = meta;
      it.off = total;
      it.q = q;
      it.w = (int64_t)pb.np[it.bp] * pb.np[it.kp];
      plan->offsets[q] = total;
      total += (int64_t)nfun(l1, B->pure[s[0]]) * nfun(l2, B->pure[s[1]]) * nfun(l3, B->pure[s[2]]) *
               nfun(l4, B->pure[s[3]]);
      buckets[cls].push_back(it);
    }
    plan->total = total;
    // flatten: classes in table order, each sorted by contraction depth
    std::vector<int> bra, ket, meta;
    std::vector<long long> off, qidx;
    bra.reserve(nq);
    ket.reserve(nq);
    meta.reserve(nq);
    off.reserve(nq);
    qidx.reserve(nq);
    int64_t ymax = 0;
    for (int cb = 0; cb < (int)buckets.size(); ++cb) {
      auto& v = buckets[cb];
      if (v.empty()) continue;
      const bool is32 = cb >= nc64;
      const int c = is32 ? cb - nc64 : cb;
      std::stable_sort(v.begin(), v.end(), [](const Item& a, const Item& b) {
        if (a.w != b.w) return a.w > b.w;
        if (a.bp != b.bp) return a.bp < b.bp;
        return a.kp < b.kp;
      });
      ClassWork cw;
      cw.cls = c;
      cw.fp32 = is32;
      cw.first = (int64_t)bra.size();
      cw.count = (int64_t)v.size();
      cw.nprimq = 0;
      for (const auto& it : v) {
        bra.push_back(it.bp);
        ket.push_back(it.kp);
        meta.push_back(it.meta);
        off.push_back(it.off);
        qidx.push_back(it.q);
        cw.nprimq += it.w;
      }
      HIPCHK(hipEventCreate(&cw.ev0));
      HIPCHK(hipEventCreate(&cw.ev1));
      plan->classes.push_back(cw);
      ymax = std::max<int64_t>(ymax, (is32 ? l2g_gen_classes32 : l2g_gen_classes)[c].ybuf_doubles);
    }
    hipStream_t st = ctx->stream;
    upload(plan->bra, bra, st);
    upload(plan->ket, ket, st);
    upload(plan->meta, meta, st);
    upload(plan->off, off, st);
    upload(plan->qidx, qidx, st);
    upload(plan->pair_first, pb.first, st);
    upload(plan->pair_np, pb.np, st);
    upload(plan->pair_AB, pb.AB, st);
    upload(plan->ppP, pb.P, st);
    upload(plan->ppPA, pb.PA, st);
    upload(plan->ppAux, pb.aux, st);
    upload(plan->ppNs, pb.ns, st);
    // persistent grid: enough blocks to fill every CU several times over
    plan->block = 128;
    const int64_t maxblocks = (int64_t)ctx->ncu * 8;
    int64_t need = 0;
    for (const auto& c : plan->classes) need = std::max(need, (c.count + plan->block - 1) / plan->block);
    plan->grid = (int)std::max<int64_t>(1, std::min(need, maxblocks));
    plan->ystride = (int64_t)plan->grid * plan->block;
    plan->ybuf.alloc(std::max<int64_t>(ymax * plan->ystride, 1) * sizeof(double));
    HIPCHK(hipStreamSynchronize(st));
    ++ctx->live;
    *out = plan.release();
  });
}

int l2g_eri_plan_create(l2g_basis B, int64_t nq, const int32_t* quartets, double precision, int screening, double scale,
                        int uniform_cart, l2g_eri_plan* out) {
  return plan_create_impl(B, nq, quartets, precision, screening, scale, uniform_cart, 0.0, out);
}

int l2g_eri_plan_create_mixed(l2g_basis B, int64_t nq, const int32_t* quartets, double precision, int screening,
                              double scale, int uniform_cart, double fp32_threshold, l2g_eri_plan* out) {
  if (!(fp32_threshold >= 0)) {
    g_err = "fp32_threshold must be >= 0";
    return L2G_E_INVALID;
  }
  return plan_create_impl(B, nq, quartets, precision, screening, scale, uniform_cart, fp32_threshold, out);
}

int l2g_eri_plan_size(l2g_eri_plan p, int64_t* total) {
  return guarded([&] {
    if (!p || !total) throw std::invalid_argument("null argument");
    *total = p->total;
  });
}

int l2g_eri_plan_offsets(l2g_eri_plan p, int64_t* offsets) {
  return guarded([&] {
    if (!p || !offsets) throw std::invalid_argument("null argument");
    std::memcpy(offsets, p->offsets.data(), p->offsets.size() * sizeof(int64_t));
  });
}

int l2g_eri_plan_execute(l2g_eri_plan p, double* d_out) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null plan");
    l2g_context_s* ctx = p->basis->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    double* out = d_out;
    if (!out) {
      p->out.alloc(std::max<int64_t>(p->total, 1) * sizeof(double));
      out = p->out.as<double>();
    }
    L2gEriArgs A;
    std::memset(&A, 0, sizeof(A));
    A.pair_first = p->pair_first.as<int>();
    A.pair_np = p->pair_np.as<int>();
    A.pair_AB = p->pair_AB.as<double>();
    A.pp_P = p->ppP.as<double4>();
    A.pp_PA = p->ppPA.as<double4>();
    A.pp_aux = p->ppAux.as<double2>();
    A.pp_nonsph = p->ppNs.as<double>();
    A.boys = ctx->boys.as<double>();
    A.precision = p->precision;
    A.ln_precision = p->ln_precision;
    A.scale = p->scale;
    A.method = p->method;
    A.shg_gaussian = ctx->shg_gaussian;
    A.uniform_cart = p->uniform;
    A.out = out;
    A.ybuf = p->ybuf.as<double>();
    A.ystride = p->ystride;
    p->nz.alloc(std::max<int64_t>(p->nq, 1));
    A.nz = p->nz.as<signed char>();
    for (auto& c : p->classes) {
      const L2gClassInfo& ci = (c.fp32 ? l2g_gen_classes32 : l2g_gen_classes)[c.cls];
      A.bra_pair = p->bra.as<int>() + c.first;
      A.ket_pair = p->ket.as<int>() + c.first;
      A.meta = p->meta.as<int>() + c.first;
      A.out_off = p->off.as<long long>() + c.first;
      A.qidx = p->qidx.as<long long>() + c.first;
      A.n = (int)c.count;
      const int grid = (int)std::min<int64_t>(p->grid, (c.count + p->block - 1) / p->block);
      HIPCHK(hipEventRecord(c.ev0, ctx->stream));
      ci.launch(A, grid, p->block, ctx->stream);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(c.ev1, ctx->stream));
    }
  });
}

int l2g_synchronize(l2g_context ctx) {
  return guarded([&] {
    if (!ctx) throw L2gError(L2G_E_STATE, "null context");
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) throw L2gError(L2G_E_LAUNCH, std::string("kernel execution failed: ") + hipGetErrorString(e));
  });
}

int l2g_eri_plan_download(l2g_eri_plan p, double* host_out, int8_t* nonzero) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null plan");
    l2g_context_s* ctx = p->basis->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) throw L2gError(L2G_E_LAUNCH, std::string("kernel execution failed: ") + hipGetErrorString(e));
    if (host_out && p->total > 0) {
      if (!p->out.p) throw L2gError(L2G_E_STATE, "plan has not been executed into its own buffer");
      HIPCHK(hipMemcpy(host_out, p->out.p, p->total * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (nonzero && p->nq > 0) {
      if (!p->nz.p) throw L2gError(L2G_E_STATE, "plan has not been executed");
      HIPCHK(hipMemcpy(nonzero, p->nz.p, p->nq, hipMemcpyDeviceToHost));
    }
  });
}

int l2g_eri_plan_device_output(l2g_eri_plan p, double** d_out) {
  return guarded([&] {
    if (!p || !d_out) throw std::invalid_argument("null argument");
    p->out.alloc(std::max<int64_t>(p->total, 1) * sizeof(double));
    *d_out = p->out.as<double>();
  });
}

int l2g_eri_plan_stats(l2g_eri_plan p, int cap, int* ncls, int* code, int64_t* nquart, int64_t* nprimq, float* ms) {
  return guarded([&] {
    if (!p || !ncls) throw std::invalid_argument("null argument");
    *ncls = (int)p->classes.size();
    for (int i = 0; i < (int)p->classes.size() && i < cap; ++i) {
      auto& c = p->classes[i];
      const auto& ci = (c.fp32 ? l2g_gen_classes32 : l2g_gen_classes)[c.cls];
      if (code) code[i] = ci.la * 1000 + ci.lb * 100 + ci.lc * 10 + ci.ld + (c.fp32 ? 10000 : 0);
      if (nquart) nquart[i] = c.count;
      if (nprimq) nprimq[i] = c.nprimq;
      if (ms) {
        float t = 0;
        if (hipEventElapsedTime(&t, c.ev0, c.ev1) != hipSuccess) t = -1;
        ms[i] = t;
      }
    }
  });
}

int l2g_eri_plan_destroy(l2g_eri_plan p) {
  if (p) {
    l2g_context_s* ctx = p->basis->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete p;
    --ctx->live;
  }
  return L2G_OK;
}

}  // extern "C"

namespace {

// Schwarz factors K of shell pairs and, when S is given, the FP32-selection
// factor of each pair in its canonical orientation,
//   S_ab = sqrt(max_e |[e0|e0]|) (1 + max_x |AB_x|)^lb,
// where [e0|e0] are the contracted pre-HRR integrals (e = la..la+lb on the
// build centre).  By the Cauchy-Schwarz inequality |[e0|f0]| <= S'_ab S'_cd,
// and the bra and ket HRR sums have absolute coefficient sums at most
// (1 + max|AB_x|)^lb and (1 + max|CD_x|)^ld, so S_ab S_cd bounds what HRR
// adds up: FP32 rounding in the recursion enters a final integral at about
// eps_fp32 S_ab S_cd.
void schwarz_impl(l2g_basis B, int64_t npairs, const int32_t* pairs, double precision, int norm, double* K,
                  double* S) {
  l2g_context_s* ctx = B->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  const double lnp = precision > 0 ? std::log(precision) : std::numeric_limits<double>::lowest();
  PairBuilder pb(*B, lnp, L2G_SCREEN_ORIGINAL);
  std::vector<std::vector<int>> pid(l2g_gen_nschwarz), meta(l2g_gen_nschwarz);
  std::vector<std::vector<long long>> idx(l2g_gen_nschwarz);
  for (int64_t q = 0; q < npairs; ++q) {
    int a = pairs[2 * q], b = pairs[2 * q + 1];
    if (a < 0 || b < 0 || a >= B->nshell || b >= B->nshell) throw std::invalid_argument("shell index out of range");
    if (B->swap_pair(a, b)) std::swap(a, b);
    const int c = ctx->schwarz_of[B->l[a]][B->l[b]];
    if (c < 0) throw L2gError(L2G_E_LMAX, "no Schwarz kernel for this shell pair");
    const int pa = (B->pure[a] && B->l[a] > 0) ? 1 : 0, pbit = (B->pure[b] && B->l[b] > 0) ? 1 : 0;
    pid[c].push_back(pb.get(a, b));
    meta[c].push_back((pa << 3) | (pbit << 4) | (pa << 5) | (pbit << 6));
    idx[c].push_back(q);
  }
  std::vector<int> allp, allm;
  std::vector<long long> alli;
  std::vector<std::pair<int, int64_t>> ranges;
  int64_t ymax = 0, maxn = 0;
  for (int c = 0; c < l2g_gen_nschwarz; ++c) {
    if (pid[c].empty()) continue;
    ranges.push_back({c, (int64_t)allp.size()});
    allp.insert(allp.end(), pid[c].begin(), pid[c].end());
    allm.insert(allm.end(), meta[c].begin(), meta[c].end());
    alli.insert(alli.end(), idx[c].begin(), idx[c].end());
    ymax = std::max<int64_t>(ymax, l2g_gen_schwarz[c].ybuf_doubles);
    maxn = std::max<int64_t>(maxn, (int64_t)pid[c].size());
  }
  hipStream_t st = ctx->stream;
  DevBuf dp, dm, di, pf, pn, pab, pP, pPA, pAux, pNs, dK, yb;
  upload(dp, allp, st);
  upload(dm, allm, st);
  upload(di, alli, st);
  upload(pf, pb.first, st);
  upload(pn, pb.np, st);
  upload(pab, pb.AB, st);
  upload(pP, pb.P, st);
  upload(pPA, pb.PA, st);
  upload(pAux, pb.aux, st);
  upload(pNs, pb.ns, st);
  dK.alloc(npairs * sizeof(double) * (S ? 2 : 1));
  const int block = 128;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((maxn + block - 1) / block, (int64_t)ctx->ncu * 8));
  const int64_t ystride = (int64_t)grid * block;
  yb.alloc(std::max<int64_t>(ymax * ystride, 1) * sizeof(double));
  L2gEriArgs A;
  std::memset(&A, 0, sizeof(A));
  A.pair_first = pf.as<int>();
  A.pair_np = pn.as<int>();
  A.pair_AB = pab.as<double>();
  A.pp_P = pP.as<double4>();
  A.pp_PA = pPA.as<double4>();
  A.pp_aux = pAux.as<double2>();
  A.pp_nonsph = pNs.as<double>();
  A.boys = ctx->boys.as<double>();
  A.precision = precision > 0 ? precision : 0.0;
  A.ln_precision = lnp;
  A.scale = 1.0;
  A.method = L2G_SCREEN_ORIGINAL;
  A.shg_gaussian = 0;
  A.schwarz_norm = norm;
  A.out = dK.as<double>();
  A.out2 = S ? dK.as<double>() + npairs : nullptr;
  A.ybuf = yb.as<double>();
  A.ystride = ystride;
  for (auto& r : ranges) {
    const int c = r.first;
    const int64_t n = (int64_t)pid[c].size();
    A.bra_pair = dp.as<int>() + r.second;
    A.ket_pair = A.bra_pair;
    A.meta = dm.as<int>() + r.second;
    A.out_off = di.as<long long>() + r.second;
    A.n = (int)n;
    l2g_gen_schwarz[c].launch(A, (int)std::min<int64_t>(grid, (n + block - 1) / block), block, st);
    HIPCHK(hipGetLastError());
  }
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) throw L2gError(L2G_E_LAUNCH, std::string("Schwarz kernel failed: ") + hipGetErrorString(e));
  HIPCHK(hipMemcpy(K, dK.p, npairs * sizeof(double), hipMemcpyDeviceToHost));
  if (S) {
    HIPCHK(hipMemcpy(S, dK.as<double>() + npairs, npairs * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < npairs; ++q) {
      const int a = pairs[2 * q], b = pairs[2 * q + 1];
      double m = 0;
      for (int k = 0; k < 3; ++k) m = std::max(m, std::fabs(B->origin[3 * a + k] - B->origin[3 * b + k]));
      S[q] *= std::pow(1.0 + m, std::min(B->l[a], B->l[b]));
    }
  }
}

}  // namespace

extern "C" {

int l2g_schwarz(l2g_basis B, int64_t npairs, const int32_t* pairs, double precision, int norm, double* K) {
  return guarded([&] {
    if (!B) throw L2gError(L2G_E_STATE, "null basis");
    if (npairs < 0 || (npairs > 0 && (!pairs || !K)) || (norm != 0 && norm != 1))
      throw std::invalid_argument("l2g_schwarz: bad arguments");
    if (npairs == 0) return;
    schwarz_impl(B, npairs, pairs, precision, norm, K, nullptr);
  });
}

int l2g_precision_bounds(l2g_basis B, int64_t npairs, const int32_t* pairs, double* K, double* S) {
  return guarded([&] {
    if (!B) throw L2gError(L2G_E_STATE, "null basis");
    if (npairs < 0 || (npairs > 0 && (!pairs || !K || !S))) throw std::invalid_argument("l2g_precision_bounds: bad arguments");
    if (npairs == 0) return;
    schwarz_impl(B, npairs, pairs, 0.0, 0, K, S);
  });
}

}  // extern "C"

namespace {

// Overlap of two contracted shells by Obara-Saika in each direction,
// S_{i+1,j} = X_PA S_ij + (i S_{i-1,j} + j S_{i,j-1}) / (2p); Frobenius norm
// of the (solid-harmonic transformed) block.  Used for the significant
// shell-pair list of compute_shellpairs (hartree-fock++.cc:1305-1436).
double overlap_norm(const l2g_basis_s& B, int a, int b) {
  const int la = B.l[a], lb = B.l[b];
  const int na = ncart(la), nb = ncart(lb);
  std::vector<double> S((size_t)na * nb, 0.0);
  double s1d[3][8][8];
  for (int pa = 0; pa < B.nprim[a]; ++pa)
    for (int pb = 0; pb < B.nprim[b]; ++pb) {
      const double al = B.alpha[B.off[a] + pa], be = B.alpha[B.off[b] + pb], p = al + be;
      const double mu = al * be / p, oo2p = 0.5 / p;
      for (int d = 0; d < 3; ++d) {
        const double A = B.origin[3 * a + d], Bc = B.origin[3 * b + d];
        const double X = A - Bc, P = (al * A + be * Bc) / p;
        const double PA = P - A, PB = P - Bc;
        s1d[d][0][0] = std::sqrt(M_PI / p) * std::exp(-mu * X * X);
        for (int i = 0; i <= la; ++i)
          for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            double v;
            if (i > 0) {
              v = PA * s1d[d][i - 1][j];
              if (i > 1) v += (i - 1) * oo2p * s1d[d][i - 2][j];
              if (j > 0) v += j * oo2p * s1d[d][i - 1][j - 1];
            } else {
              v = PB * s1d[d][i][j - 1];
              if (j > 1) v += (j - 1) * oo2p * s1d[d][i][j - 2];
            }
            s1d[d][i][j] = v;
          }
      }
      const double c = B.coeff[B.off[a] + pa] * B.coeff[B.off[b] + pb];
      int ia = 0;
      for (int x1 = la; x1 >= 0; --x1)
        for (int y1 = la - x1; y1 >= 0; --y1, ++ia) {
          const int z1 = la - x1 - y1;
          int ib = 0;
          for (int x2 = lb; x2 >= 0; --x2)
            for (int y2 = lb - x2; y2 >= 0; --y2, ++ib) {
              const int z2 = lb - x2 - y2;
              S[(size_t)ia * nb + ib] += c * s1d[0][x1][x2] * s1d[1][y1][y2] * s1d[2][z1][z2];
            }
        }
    }
  auto sph = [&](int l, std::vector<double>& M) {
    M.assign((size_t)(2 * l + 1) * ncart(l), 0.0);
    for (int m = -l; m <= l; ++m) {
      int ic = 0;
      for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y, ++ic) M[(size_t)(m + l) * ncart(l) + ic] = l2g_solid_harmonics_coeff(l, m, x, y, l - x - y);
    }
  };
  int n1 = na, n2 = nb;
  std::vector<double> cur = S, M;
  if (B.pure[a] && la > 0) {
    sph(la, M);
    std::vector<double> nx((size_t)(2 * la + 1) * n2, 0.0);
    for (int r = 0; r < 2 * la + 1; ++r)
      for (int c = 0; c < n1; ++c)
        for (int j = 0; j < n2; ++j) nx[(size_t)r * n2 + j] += M[(size_t)r * n1 + c] * cur[(size_t)c * n2 + j];
    cur.swap(nx);
    n1 = 2 * la + 1;
  }
  if (B.pure[b] && lb > 0) {
    sph(lb, M);
    std::vector<double> nx((size_t)n1 * (2 * lb + 1), 0.0);
    for (int i = 0; i < n1; ++i)
      for (int r = 0; r < 2 * lb + 1; ++r)
        for (int c = 0; c < n2; ++c) nx[(size_t)i * (2 * lb + 1) + r] += M[(size_t)r * n2 + c] * cur[(size_t)i * n2 + c];
    cur.swap(nx);
  }
  double nrm = 0;
  for (double v : cur) nrm += v * v;
  return std::sqrt(nrm);
}

template <class F>
void parallel_for(int n, F&& f) {
  const int nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  std::atomic<int> next{0};
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (int i = next++; i < n; i = next++) f(i);
    });
  for (auto& t : th) t.join();
}

// launches one generated kernel family over per-class work lists
struct WorkLists {
  std::vector<int> bra, ket, meta;
  std::vector<long long> off;
  std::vector<double> lnp, scale;
  std::vector<int4> bf;
};

}  // namespace

struct l2g_fock_s {
  l2g_basis_s* basis;
  int method;
  double max_engine_precision;
  int nbf = 0;
  std::vector<int> bf1st;
  std::vector<std::vector<int>> splist;   // splist[s1] ascending s2 <= s1
  std::vector<double> K;                  // nsh x nsh Schwarz (inf norm), significant pairs only
  std::vector<double> S;                  // nsh x nsh FP32-selection factors (schwarz_impl)
  std::vector<std::vector<double>> pfac;  // primitive Schwarz factors per significant pair (loop order)
  std::vector<uint64_t> pfac_key;
  int64_t last_nq = 0, last_nprim = 0, last_nq32 = 0;
  float last_ms = 0;
  double fp32_threshold = 0;  // quartets with K12 K34 below it are evaluated in FP32
  DevBuf dens, jg, kg;
};

extern "C" {

int l2g_fock_create(l2g_basis B, double pair_threshold, double max_engine_precision, int screening, l2g_fock* out) {
  return guarded([&] {
    if (!B || !out) throw std::invalid_argument("l2g_fock_create: null argument");
    if (screening != L2G_SCREEN_ORIGINAL && screening != L2G_SCREEN_CONSERVATIVE && screening != L2G_SCREEN_SCHWARZ &&
        screening != L2G_SCREEN_SCHWARZINF)
      throw std::invalid_argument("unknown screening method");
    auto f = std::make_unique<l2g_fock_s>();
    f->basis = B;
    f->method = screening;
    f->max_engine_precision = max_engine_precision;
    const int n = B->nshell;
    f->bf1st.resize(n + 1);
    f->bf1st[0] = 0;
    for (int s = 0; s < n; ++s) f->bf1st[s + 1] = f->bf1st[s] + nfun(B->l[s], B->pure[s]);
    f->nbf = f->bf1st[n];
    // significant shell pairs: same centre, or overlap-block norm >= threshold
    f->splist.assign(n, {});
    parallel_for(n, [&](int s1) {
      for (int s2 = 0; s2 <= s1; ++s2) {
        const bool same = B->origin[3 * s1] == B->origin[3 * s2] && B->origin[3 * s1 + 1] == B->origin[3 * s2 + 1] &&
                          B->origin[3 * s1 + 2] == B->origin[3 * s2 + 2];
        if (same || overlap_norm(*B, s1, s2) >= pair_threshold) f->splist[s1].push_back(s2);
      }
    });
    std::vector<int32_t> pairs;
    for (int s1 = 0; s1 < n; ++s1)
      for (int s2 : f->splist[s1]) {
        pairs.push_back(s1);
        pairs.push_back(s2);
      }
    const int64_t np = (int64_t)pairs.size() / 2;
    // Schwarz factors of the significant pairs, no primitive screening
    std::vector<double> Kp(np), Sp(np);
    if (np > 0) schwarz_impl(B, np, pairs.data(), 0.0, 0, Kp.data(), Sp.data());
    f->K.assign((size_t)n * n, 0.0);
    f->S.assign((size_t)n * n, 0.0);
    for (int64_t i = 0; i < np; ++i) {
      const int a = pairs[2 * i], b = pairs[2 * i + 1];
      f->K[(size_t)a * n + b] = f->K[(size_t)b * n + a] = Kp[i];
      f->S[(size_t)a * n + b] = f->S[(size_t)b * n + a] = Sp[i];
    }
    int rc = 0;
    // primitive-pair Schwarz factors from normalization-free primitives
    // (Shell::extract_primitive(p, false), shell.h:938-947), on the GPU
    if (screening == L2G_SCREEN_SCHWARZ || screening == L2G_SCREEN_SCHWARZINF) {
      std::vector<int> pl, pp, pn;
      std::vector<double> pa, pc, po;
      std::vector<int> first(n);
      for (int s = 0; s < n; ++s) {
        first[s] = (int)pl.size();
        for (int k = 0; k < B->nprim[s]; ++k) {
          pl.push_back(B->l[s]);
          pp.push_back(B->pure[s]);
          pn.push_back(1);
          pa.push_back(B->alpha[B->off[s] + k]);
          pc.push_back(1.0);
          for (int d = 0; d < 3; ++d) po.push_back(B->origin[3 * s + d]);
        }
      }
      l2g_basis PB = nullptr;
      rc = l2g_basis_create(B->ctx, (int)pl.size(), pl.data(), pp.data(), pn.data(), pa.data(), pc.data(), po.data(), &PB);
      if (rc) throw L2gError(rc, g_err);
      std::unique_ptr<l2g_basis_s> hold(PB);
      std::vector<int32_t> ppairs;
      for (int64_t i = 0; i < np; ++i) {
        const int a = pairs[2 * i], b = pairs[2 * i + 1];
        for (int p1 = 0; p1 < B->nprim[a]; ++p1)
          for (int p2 = 0; p2 < B->nprim[b]; ++p2) {
            ppairs.push_back(first[a] + p1);
            ppairs.push_back(first[b] + p2);
          }
      }
      std::vector<double> kf(ppairs.size() / 2);
      rc = l2g_schwarz(PB, (int64_t)kf.size(), ppairs.data(), 0.0, screening == L2G_SCREEN_SCHWARZINF ? 0 : 1, kf.data());
      if (rc) throw L2gError(rc, g_err);
      f->pfac.resize(np);
      f->pfac_key.resize(np);
      size_t k = 0;
      for (int64_t i = 0; i < np; ++i) {
        const int a = pairs[2 * i], b = pairs[2 * i + 1];
        f->pfac[i].assign(kf.begin() + k, kf.begin() + k + (size_t)B->nprim[a] * B->nprim[b]);
        k += (size_t)B->nprim[a] * B->nprim[b];
        f->pfac_key[i] = (uint64_t)(uint32_t)a << 32 | (uint32_t)b;
      }
    }
    ++B->ctx->live;
    *out = f.release();
  });
}

int l2g_fock_nbf(l2g_fock f, int* nbf) {
  return guarded([&] {
    if (!f || !nbf) throw std::invalid_argument("null argument");
    *nbf = f->nbf;
  });
}

int l2g_fock_schwarz(l2g_fock f, double* K) {
  return guarded([&] {
    if (!f || !K) throw std::invalid_argument("null argument");
    std::memcpy(K, f->K.data(), f->K.size() * sizeof(double));
  });
}

int l2g_fock_build(l2g_fock f, const double* D, double precision, double* J, double* K, double* G) {
  return guarded([&] {
    if (!f || !D) throw std::invalid_argument("l2g_fock_build: null argument");
    l2g_basis_s* B = f->basis;
    l2g_context_s* ctx = B->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = B->nshell, N = f->nbf;
    // shell-block infinity norms of D (compute_shellblock_norm)
    std::vector<double> Dn((size_t)n * n, 0.0);
    double dmax = 0;
    for (int s1 = 0; s1 < n; ++s1)
      for (int s2 = 0; s2 < n; ++s2) {
        double m = 0;
        for (int i = f->bf1st[s1]; i < f->bf1st[s1 + 1]; ++i)
          for (int j = f->bf1st[s2]; j < f->bf1st[s2 + 1]; ++j) m = std::max(m, std::fabs(D[(size_t)i * N + j]));
        Dn[(size_t)s1 * n + s2] = m;
        dmax = std::max(dmax, m);
      }
    const double needed = dmax > 0 ? precision / dmax : precision;
    if (needed < f->max_engine_precision)
      throw std::invalid_argument("requested Fock precision needs more engine precision than the shell-pair data hold");
    PairBuilder pb(*B, std::log(f->max_engine_precision), f->method);
    for (size_t i = 0; i < f->pfac.size(); ++i) pb.pfac[f->pfac_key[i]] = &f->pfac[i];
    // permutationally-unique screened quartets (hartree-fock++.cc:1628-1696)
    const int nf64 = l2g_gen_nfock;
    std::vector<WorkLists> W(l2g_gen_nfock + l2g_gen_nfock32);
    int64_t nq32 = 0;
    int64_t nq = 0, nprim = 0;
    for (int s1 = 0; s1 < n; ++s1)
      for (int s2 : f->splist[s1]) {
        const double Dn12 = Dn[(size_t)s1 * n + s2];
        for (int s3 = 0; s3 <= s1; ++s3) {
          const double Dn123 = std::max(std::max(Dn[(size_t)s1 * n + s3], Dn[(size_t)s2 * n + s3]), Dn12);
          const int s4max = (s1 == s3) ? s2 : s3;
          for (int s4 : f->splist[s3]) {
            if (s4 > s4max) break;
            const double Dn1234 = std::max(std::max(Dn[(size_t)s1 * n + s4], Dn[(size_t)s2 * n + s4]),
                                           std::max(Dn[(size_t)s3 * n + s4], Dn123));
            if (Dn1234 * f->K[(size_t)s1 * n + s2] * f->K[(size_t)s3 * n + s4] < precision) continue;
            const double qprec = Dn1234 != 0. ? precision / Dn1234 : needed;
            const int deg = (s1 == s2 ? 1 : 2) * (s3 == s4 ? 1 : 2) * (s1 == s3 ? (s2 == s4 ? 1 : 2) : 2);
            // canonical order (engine.impl.h:1181-1208), equal-l pairs
            // oriented by second_is_centre
            const int l1 = B->l[s1], l2 = B->l[s2], l3 = B->l[s3], l4 = B->l[s4];
            const bool stb = B->swap_pair(s1, s2), stk = B->swap_pair(s3, s4), sbk = (l1 + l2) > (l3 + l4);
            const int b1 = sbk ? (stk ? s4 : s3) : (stb ? s2 : s1);
            const int b2 = sbk ? (stk ? s3 : s4) : (stb ? s1 : s2);
            const int k1 = sbk ? (stb ? s2 : s1) : (stk ? s4 : s3);
            const int k2 = sbk ? (stb ? s1 : s2) : (stk ? s3 : s4);
            int cls = ctx->fock_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
            if (cls < 0) throw L2gError(L2G_E_LMAX, "no Fock kernel for this angular momentum class");
            if (f->fp32_threshold > 0 && f->K[(size_t)s1 * n + s2] * f->K[(size_t)s3 * n + s4] < f->fp32_threshold &&
                f->S[(size_t)s1 * n + s2] * f->S[(size_t)s3 * n + s4] < L2G_FP32_HRR_FACTOR * f->fp32_threshold) {
              const int c32 = ctx->fock32_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
              if (c32 >= 0) {
                cls = nf64 + c32;
                ++nq32;
              }
            }
            auto purebit = [&](int sh) { return (B->pure[sh] && B->l[sh] > 0) ? 1 : 0; };
            WorkLists& w = W[cls];
            const int bp = pb.get(b1, b2), kp = pb.get(k1, k2);
            w.bra.push_back(bp);
            w.ket.push_back(kp);
            w.meta.push_back((purebit(b1) << 3) | (purebit(b2) << 4) | (purebit(k1) << 5) | (purebit(k2) << 6));
            w.off.push_back(0);
            w.lnp.push_back(qprec > 0 ? std::log(qprec) : std::numeric_limits<double>::lowest());
            w.scale.push_back((double)deg);
            w.bf.push_back(int4{f->bf1st[b1], f->bf1st[b2], f->bf1st[k1], f->bf1st[k2]});
            ++nq;
            nprim += (int64_t)pb.np[bp] * pb.np[kp];
          }
        }
      }
    f->last_nq = nq;
    f->last_nprim = nprim;
    f->last_nq32 = nq32;
    auto finfo = [&](int c) -> const L2gClassInfo& { return c < nf64 ? l2g_gen_fock[c] : l2g_gen_fock32[c - nf64]; };
    hipStream_t st = ctx->stream;
    // device data
    DevBuf pf, pn, pab, pP, pPA, pAux, pNs, yb;
    upload(pf, pb.first, st);
    upload(pn, pb.np, st);
    upload(pab, pb.AB, st);
    upload(pP, pb.P, st);
    upload(pPA, pb.PA, st);
    upload(pAux, pb.aux, st);
    upload(pNs, pb.ns, st);
    const size_t NN = (size_t)N * N;
    f->dens.alloc(NN * sizeof(double));
    f->jg.alloc(NN * sizeof(double));
    f->kg.alloc(NN * sizeof(double));
    HIPCHK(hipMemcpyAsync(f->dens.p, D, NN * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(f->jg.p, 0, NN * sizeof(double), st));
    HIPCHK(hipMemsetAsync(f->kg.p, 0, NN * sizeof(double), st));
    const int block = 128;
    int64_t maxn = 0, ymax = 0;
    for (int c = 0; c < (int)W.size(); ++c) {
      maxn = std::max<int64_t>(maxn, (int64_t)W[c].bra.size());
      if (!W[c].bra.empty()) ymax = std::max<int64_t>(ymax, finfo(c).ybuf_doubles);
    }
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((maxn + block - 1) / block, (int64_t)ctx->ncu * 8));
    const int64_t ystride = (int64_t)grid * block;
    yb.alloc(std::max<int64_t>(ymax * ystride, 1) * sizeof(double));
    std::vector<std::unique_ptr<DevBuf[]>> keep;
    std::deque<WorkLists> hostkeep;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    for (int c = 0; c < (int)W.size(); ++c) {
      WorkLists& w = W[c];
      if (w.bra.empty()) continue;
      // sort by contraction depth for SIMT uniformity
      const int64_t m = (int64_t)w.bra.size();
      std::vector<int64_t> ord(m);
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
        return (int64_t)pb.np[w.bra[a]] * pb.np[w.ket[a]] > (int64_t)pb.np[w.bra[b]] * pb.np[w.ket[b]];
      });
      // host arrays must outlive the asynchronous uploads: keep them until
      // the stream is synchronised below
      hostkeep.emplace_back();
      WorkLists& s = hostkeep.back();
      for (int64_t i : ord) {
        s.bra.push_back(w.bra[i]);
        s.ket.push_back(w.ket[i]);
        s.meta.push_back(w.meta[i]);
        s.off.push_back(0);
        s.lnp.push_back(w.lnp[i]);
        s.scale.push_back(w.scale[i]);
        s.bf.push_back(w.bf[i]);
      }
      keep.emplace_back(new DevBuf[7]);
      DevBuf* d = keep.back().get();
      upload(d[0], s.bra, st);
      upload(d[1], s.ket, st);
      upload(d[2], s.meta, st);
      upload(d[3], s.off, st);
      upload(d[4], s.lnp, st);
      upload(d[5], s.scale, st);
      upload(d[6], s.bf, st);
      L2gEriArgs A;
      std::memset(&A, 0, sizeof(A));
      A.bra_pair = d[0].as<int>();
      A.ket_pair = d[1].as<int>();
      A.meta = d[2].as<int>();
      A.out_off = d[3].as<long long>();
      A.n = (int)m;
      A.pair_first = pf.as<int>();
      A.pair_np = pn.as<int>();
      A.pair_AB = pab.as<double>();
      A.pp_P = pP.as<double4>();
      A.pp_PA = pPA.as<double4>();
      A.pp_aux = pAux.as<double2>();
      A.pp_nonsph = pNs.as<double>();
      A.boys = ctx->boys.as<double>();
      A.scale = 1.0;
      A.method = f->method;
      A.shg_gaussian = ctx->shg_gaussian;
      A.ybuf = yb.as<double>();
      A.ystride = ystride;
      A.q_lnprec = d[4].as<double>();
      A.q_scale = d[5].as<double>();
      A.q_bf = d[6].as<int4>();
      A.dens = f->dens.as<double>();
      A.jg = f->jg.as<double>();
      A.kg = f->kg.as<double>();
      A.nbf = N;
      finfo(c).launch(A, (int)std::min<int64_t>(grid, (m + block - 1) / block), block, st);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(e1, st));
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) throw L2gError(L2G_E_LAUNCH, std::string("Fock kernels failed: ") + hipGetErrorString(e));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    f->last_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<double> jg(NN), kg(NN);
    HIPCHK(hipMemcpy(jg.data(), f->jg.p, NN * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(kg.data(), f->kg.p, NN * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        const double Jv = 0.25 * (jg[(size_t)i * N + j] + jg[(size_t)j * N + i]);
        const double Kv = 0.125 * (kg[(size_t)i * N + j] + kg[(size_t)j * N + i]);
        if (J) J[(size_t)i * N + j] = Jv;
        if (K) K[(size_t)i * N + j] = Kv;
        if (G) G[(size_t)i * N + j] = 2 * Jv - Kv;
      }
  });
}

int l2g_fock_set_fp32_threshold(l2g_fock f, double threshold) {
  return guarded([&] {
    if (!f) throw std::invalid_argument("null argument");
    if (!(threshold >= 0)) throw std::invalid_argument("threshold must be >= 0");
    f->fp32_threshold = threshold;
  });
}

int l2g_fock_stats_mixed(l2g_fock f, int64_t* nquartets_fp32) {
  return guarded([&] {
    if (!f) throw std::invalid_argument("null argument");
    if (nquartets_fp32) *nquartets_fp32 = f->last_nq32;
  });
}

int l2g_fock_stats(l2g_fock f, int64_t* nquartets, int64_t* nprim_quartets, float* kernel_ms) {
  return guarded([&] {
    if (!f) throw std::invalid_argument("null argument");
    if (nquartets) *nquartets = f->last_nq;
    if (nprim_quartets) *nprim_quartets = f->last_nprim;
    if (kernel_ms) *kernel_ms = f->last_ms;
  });
}

int l2g_fock_destroy(l2g_fock f) {
  if (f) {
    l2g_context_s* ctx = f->basis->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete f;
    --ctx->live;
  }
  return L2G_OK;
}

int l2g_eri_compute(l2g_basis B, int64_t nq, const int32_t* quartets, double precision, int screening, double* host_out,
                    int64_t* offsets, int8_t* nonzero) {
  l2g_eri_plan p = nullptr;
  int rc = l2g_eri_plan_create(B, nq, quartets, precision, screening, 1.0, 0, &p);
  if (rc) return rc;
  if (offsets) rc = l2g_eri_plan_offsets(p, offsets);
  if (!rc) rc = l2g_eri_plan_execute(p, nullptr);
  if (!rc) rc = l2g_eri_plan_download(p, host_out, nonzero);
  l2g_eri_plan_destroy(p);
  return rc;
}

}  // extern "C"
