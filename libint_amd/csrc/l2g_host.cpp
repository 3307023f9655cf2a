// Host side of the C ABI declared in include/l2g.h: device context, Boys
// table, basis upload, shell-pair data, batch planning, Schwarz factors and
// kernel launches.  The Fock build and the request server are in l2g_fock.cpp.
#include <cstdio>
#include <mutex>
#include <numeric>

#include "l2g_internal.h"

using namespace l2g_impl;

namespace {

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

}  // namespace

struct ClassWork {
  int cls;  // index into l2g_gen_classes (l2g_gen_classes32 when fp32, l2g_gen_generic when generic)
  int spec_cls = -1;            // the class's own table index (cls before the layout's choice)
  bool fp32 = false;
  bool generic = false;
  bool native_generic = false;  // no register kernel for this class
  bool coop = false;  // a wave per quartet (few, deeply contracted quartets)
  int gidx = -1;      // the class's program in the generic kernel's tables
  int64_t first, count;
  ClassSlot slot;     // grid, scratch region and stream of the launch
  int64_t nprimq;
  // generic classes of few deep quartets: the quartets split into pieces of
  // at most `chunk` primitive quartets, a block each, then summed per quartet
  int64_t npieces = 0, piece_off = 0, pfirst_off = 0, gpart_off = 0, gstride = 0;
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
  std::vector<int> itemw;  // primitive quartets of each work item (host)
  DevBuf gp_bra, gp_ket, gp_meta, gp_qidx, gp_range, gp_first, gpart;  // generic-class pieces
  int64_t ymax = 0;  // largest scratch doubles per lane over the classes
  int block = 128;
  bool wave_split = true;  // classes of few deep quartets run a wave per quartet
  bool few_generic = true;  // ... or the generic kernel, where measured faster
  int oper = 0;             // 0 Coulomb, 1 erf-attenuated, 2 erfc-attenuated
  double omega = 0;
  // hipGraph mode: the class launches are captured once per output buffer
  // and replayed; per-class timings are then not recorded
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double* graph_out = nullptr;
  int graph_gauss = -1;  // solid-harmonic ordering the graph's launches were captured with
  hipEvent_t t0 = nullptr, t1 = nullptr;  // whole execution
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr, graph = nullptr, graph_out = nullptr;
    graph_gauss = -1;
  }
  ~l2g_eri_plan_s() {
    drop_graph();
    for (auto& c : classes) {
      if (c.ev0) (void)hipEventDestroy(c.ev0);
      if (c.ev1) (void)hipEventDestroy(c.ev1);
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
  }
};

namespace {
void plan_layout(l2g_eri_plan_s* p, l2g_context_s* ctx);  // launch layout, defined below
}  // namespace


// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* l2g_version(void) { return "libint_amd 0.1 (gfx950)"; }

const char* l2g_last_error(void) { return g_err.c_str(); }

int l2g_max_l(void) { return lmax_supported(); }

int l2g_fp32_max_l(void) {
  int m = -1;
  for (int i = 0; i < l2g_gen_nclasses32; ++i) {
    const L2gClassInfo& c = l2g_gen_classes32[i];
    if (c.launch) m = std::max(m, std::max(std::max(c.la, c.lb), std::max(c.lc, c.ld)));
  }
  return m;
}

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

extern "C++" void l2g_impl::register_exit_handler() {
  static std::once_flag once;
  std::call_once(once, [] {
    std::atexit([] {
      l2g_impl::stop_servers_at_exit();
      l2g_impl::g_exiting = true;
    });
  });
}

int l2g_init(int device, l2g_context* out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("l2g_init: null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw L2gError(L2G_E_NODEVICE, "no HIP device available");
    if (device < 0 || device >= n) throw std::invalid_argument("l2g_init: device index out of range");
    register_exit_handler();  // after the HIP runtime has initialised, so it runs before its teardown
    auto ctx = std::make_unique<l2g_context_s>();
    ctx->device = device;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    ctx->ncu = prop.multiProcessorCount;
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (int k = 0; k < l2g_context_s::kSide; ++k) {
      HIPCHK(hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&ctx->join[k], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
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
    fill(ctx->classerf_of, l2g_gen_classes_erf, l2g_gen_nclasses_erf);
    fill(ctx->fock32_of, l2g_gen_fock32, l2g_gen_nfock32);
    for (auto& a : ctx->schwarz_of)
      for (auto& b : a) b = -1;
    for (int i = 0; i < l2g_gen_nschwarz; ++i) {
      const auto& c = l2g_gen_schwarz[i];
      if (c.launch) ctx->schwarz_of[c.la][c.lb] = i;
    }
    for (auto& a : ctx->generic_of)
      for (auto& b : a)
        for (auto& c : b)
          for (auto& d : c) d = -1;
    for (int i = 0; i < l2g_gen_ngeneric; ++i) {
      const auto& c = l2g_gen_generic[i];
      ctx->generic_of[c.la][c.lb][c.lc][c.ld] = i;
    }
    ctx->gprog.alloc(std::max<size_t>((size_t)l2g_gen_generic_nprog, 1) * sizeof(int4));
    ctx->gidx.alloc(std::max<size_t>((size_t)l2g_gen_generic_nidx, 1) * sizeof(int));
    if (l2g_gen_generic_nprog)
      HIPCHK(hipMemcpy(ctx->gprog.p, l2g_gen_generic_prog, (size_t)l2g_gen_generic_nprog * sizeof(int4),
                       hipMemcpyHostToDevice));
    if (l2g_gen_generic_nidx)
      HIPCHK(hipMemcpy(ctx->gidx.p, l2g_gen_generic_idx, (size_t)l2g_gen_generic_nidx * sizeof(int),
                       hipMemcpyHostToDevice));
    // solid-harmonic coefficients of the generic kernel, standard order
    std::vector<double> sh(4 * 7 * 10, 0.0);
    for (int l = 0; l <= 3; ++l)
      for (int m = -l; m <= l; ++m) {
        int c = 0;
        for (int x = l; x >= 0; --x)
          for (int y = l - x; y >= 0; --y, ++c) sh[(l * 7 + m + l) * 10 + c] = l2g_solid_harmonics_coeff(l, m, x, y, l - x - y);
      }
    generic_init(sh.data());
    noalloc_setup(device);  // this device's address of the no-allocation scratch array
    *out = ctx.release();
  });
}

int l2g_fp32_paying_classes(int cap, int* n, int* codes) {
  return guarded([&] {
    if (!n || (cap > 0 && !codes)) throw std::invalid_argument("l2g_fp32_paying_classes: null argument");
    int k = 0;
    for (int i = 0; i < l2g_gen_nclasses32; ++i) {
      const L2gClassInfo& c = l2g_gen_classes32[i];
      if (!c.launch || !c.fp32_pays) continue;
      if (k < cap) codes[k] = c.la * 1000 + c.lb * 100 + c.lc * 10 + c.ld;
      ++k;
    }
    *n = k;
  });
}

int l2g_set_fp32_policy(l2g_context ctx, int policy) {
  return guarded([&] {
    if (!ctx) throw L2gError(L2G_E_STATE, "null context");
    if (policy != L2G_FP32_MEASURED && policy != L2G_FP32_ALWAYS)
      throw std::invalid_argument("l2g_set_fp32_policy: unknown policy");
    ctx->fp32_policy = policy;
  });
}

int l2g_finalize(l2g_context ctx) {
  return guarded([&] {
    if (!ctx || g_exiting) return;
    if (ctx->live > 0)
      throw L2gError(L2G_E_STATE, std::to_string(ctx->live.load()) +
                                      " basis/plan/Fock object(s) of this context are still alive; destroy them first");
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int k = 0; k < l2g_context_s::kSide; ++k) {
      if (ctx->side[k]) (void)hipStreamSynchronize(ctx->side[k]), (void)hipStreamDestroy(ctx->side[k]);
      if (ctx->join[k]) (void)hipEventDestroy(ctx->join[k]);
    }
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int l2g_set_generic_kernels(l2g_context ctx, int on) {
  return guarded([&] {
    if (!ctx) throw L2gError(L2G_E_STATE, "null context");
    ctx->force_generic = on != 0;
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
      if (l[s] > lmax_supported())
        throw L2gError(L2G_E_LMAX, "angular momentum " + std::to_string(l[s]) + " exceeds the compiled limit " +
                                       std::to_string(lmax_supported()));
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
  if (!g_exiting) basis_release(b);
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
    const int nc32 = l2g_gen_nclasses32;
    std::vector<std::vector<Item>> buckets(l2g_gen_nclasses + l2g_gen_nclasses32 + l2g_gen_ngeneric);
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
      int cls = ctx->force_generic ? -1 : ctx->class_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
      if (cls < 0) {
        const int g = ctx->generic_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
        if (g < 0) throw L2gError(L2G_E_LMAX, "no kernel for this angular momentum class");
        cls = nc64 + nc32 + g;  // FP64 only
      } else if (fp32_threshold > 0 && use_fp32(s)) {
        const int c32 = ctx->class32_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
        if (c32 >= 0 && (ctx->fp32_policy || l2g_gen_classes32[c32].fp32_pays)) cls = nc64 + c32;
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
    // measured routing: an FP32 launch too small to fill the chip (its
    // lanes waiting on a tail) costs more than FP32 saves, so below
    // 4 x 256 lanes per CU a class's FP32 items run in its FP64 launch
    if (!ctx->fp32_policy) {
      const size_t min_items = (size_t)4 * ctx->ncu * 256;
      for (int c32 = 0; c32 < nc32; ++c32) {
        auto& v = buckets[nc64 + c32];
        if (v.empty() || v.size() >= min_items) continue;
        const L2gClassInfo& ci = l2g_gen_classes32[c32];
        const int c64 = ctx->class_of[ci.la][ci.lb][ci.lc][ci.ld];
        if (c64 < 0) continue;
        auto& d = buckets[c64];
        d.insert(d.end(), v.begin(), v.end());
        v.clear();
      }
    }
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
      const bool isgen = cb >= nc64 + nc32;
      const bool is32 = !isgen && cb >= nc64;
      const int c = isgen ? cb - nc64 - nc32 : (is32 ? cb - nc64 : cb);
      std::stable_sort(v.begin(), v.end(), [](const Item& a, const Item& b) {
        if (a.w != b.w) return a.w > b.w;
        if (a.bp != b.bp) return a.bp < b.bp;
        return a.kp < b.kp;
      });
      ClassWork cw;
      cw.cls = c;
      cw.spec_cls = c;
      cw.fp32 = is32;
      cw.generic = isgen;
      cw.native_generic = isgen;
      {
        const L2gClassInfo* ci = isgen ? nullptr : &(is32 ? l2g_gen_classes32 : l2g_gen_classes)[c];
        cw.gidx = isgen ? c : ctx->generic_of[ci->la][ci->lb][ci->lc][ci->ld];
      }
      cw.first = (int64_t)bra.size();
      cw.count = (int64_t)v.size();
      cw.nprimq = 0;
      for (const auto& it : v) {
        bra.push_back(it.bp);
        ket.push_back(it.kp);
        meta.push_back(it.meta);
        off.push_back(it.off);
        qidx.push_back(it.q);
        plan->itemw.push_back(it.w);
        cw.nprimq += it.w;
      }
      HIPCHK(hipEventCreate(&cw.ev0));
      HIPCHK(hipEventCreate(&cw.ev1));
      plan->classes.push_back(cw);
      if (!isgen) ymax = std::max<int64_t>(ymax, (is32 ? l2g_gen_classes32 : l2g_gen_classes)[c].ybuf_doubles);
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
    plan->ymax = ymax;
    plan_layout(plan.get(), ctx);
    HIPCHK(hipStreamSynchronize(st));
    ++ctx->live;
    basis_retain(B);
    *out = plan.release();
  });
}

}  // extern "C"

namespace {

// Launch layout of a plan: a persistent grid with enough blocks to fill
// every CU several times over, and the lane-private scratch for it.  A class
// of few quartets (fewer than a quarter of the chip's lanes at one wave per
// SIMD) with deep primitive loops would, at one lane per quartet, leave most
// of the chip idle while a few lanes run their loops serially.  With
// wave_split (the default) such a class runs a wave per quartet whose lanes
// split the loop (summing each contraction as 64 serial partial sums and a
// fixed butterfly, another order than libint's serial loop, no less
// accurate); without it, a class of total angular momentum 7 or more averaging at least
// 64 primitive quartets per quartet runs on the generic kernel, a block per
// quartet whose threads split the targets and keep each contraction's
// serial order.  Lower classes have too few targets per quartet to occupy
// a block and stay at a lane per quartet (on 400k quartets of an s-f basis
// with K up to 8, the cut at L >= 7 measured 180 ms, at 5, 6, 8 339, 276
// and 207 ms, and 1556 ms without the generic kernel).
#ifndef L2G_GENERIC_MIN_DEPTH
#define L2G_GENERIC_MIN_DEPTH 64
#endif
#ifndef L2G_GENERIC_MIN_L
#define L2G_GENERIC_MIN_L 7
#endif

// total angular momentum of a class with a register kernel
int class_l(const ClassWork& c) {
  if (c.native_generic) return 99;
  const L2gClassInfo& ci = (c.fp32 ? l2g_gen_classes32 : l2g_gen_classes)[c.spec_cls];
  return ci.la + ci.lb + ci.lc + ci.ld;
}

// Generic-kernel classes of few quartets (up to 16 per CU) whose deepest
// quartet has more primitive quartets than a piece of about 1/(4 ncu) of
// the class's (at least 32): each quartet splits into pieces of at most
// that many, a block each, that leave their contracted [e0|f0] values in
// gpart; a second pass sums each quartet's pieces in piece order and runs
// the HRR and store.  A block per quartet would otherwise walk the deepest
// quartet's primitive loop alone while the rest of the chip idles.
void plan_pieces(l2g_eri_plan_s* p, l2g_context_s* ctx) {
  std::vector<int> pb, pk, pm, pf;
  std::vector<long long> pq;
  std::vector<int2> pr;
  std::vector<int> hb((size_t)p->nq), hk((size_t)p->nq), hm((size_t)p->nq);
  std::vector<long long> hq((size_t)p->nq);
  bool have = false;
  int64_t gtotal = 0;
  for (auto& c : p->classes) {
    c.npieces = 0;
    if (!c.generic || c.count > (int64_t)ctx->ncu * 16) continue;
    int wmax = 0;
    for (int64_t i = c.first; i < c.first + c.count; ++i) wmax = std::max(wmax, p->itemw[(size_t)i]);
    int64_t chunk = std::max<int64_t>(32, (c.nprimq + 4 * ctx->ncu - 1) / (4 * ctx->ncu));
    chunk += chunk & 1;  // even: the two-primitive pass stays full
    if (wmax <= chunk) continue;
    if (!have) {  // the plan's work-item arrays, read back once
      HIPCHK(hipMemcpy(hb.data(), p->bra.p, hb.size() * sizeof(int), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(hk.data(), p->ket.p, hk.size() * sizeof(int), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(hm.data(), p->meta.p, hm.size() * sizeof(int), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(hq.data(), p->qidx.p, hq.size() * sizeof(long long), hipMemcpyDeviceToHost));
      have = true;
    }
    const L2gGenericClass& g = l2g_gen_generic[c.cls];
    c.gstride = (int64_t)g.ne * g.accs + 1;
    c.piece_off = (int64_t)pb.size();
    c.pfirst_off = (int64_t)pf.size();
    for (int64_t i = c.first; i < c.first + c.count; ++i) {
      pf.push_back((int)((int64_t)pb.size() - c.piece_off));
      const int w = p->itemw[(size_t)i];
      for (int a = 0; a < w || a == 0; a += (int)chunk) {
        pb.push_back(hb[(size_t)i]);
        pk.push_back(hk[(size_t)i]);
        pm.push_back(hm[(size_t)i]);
        pq.push_back(hq[(size_t)i]);
        pr.push_back(make_int2(a, (int)std::min<int64_t>(w, a + chunk)));
      }
    }
    pf.push_back((int)((int64_t)pb.size() - c.piece_off));
    c.npieces = (int64_t)pb.size() - c.piece_off;
    c.gpart_off = gtotal;
    gtotal += c.npieces * c.gstride;
  }
  if (pb.empty()) return;
  hipStream_t st = ctx->stream;
  upload(p->gp_bra, pb, st);
  upload(p->gp_ket, pk, st);
  upload(p->gp_meta, pm, st);
  upload(p->gp_qidx, pq, st);
  upload(p->gp_range, pr, st);
  upload(p->gp_first, pf, st);
  p->gpart.alloc(gtotal * sizeof(double));
  HIPCHK(hipStreamSynchronize(st));
}

void plan_layout(l2g_eri_plan_s* p, l2g_context_s* ctx) {
  p->block = 128;
  std::vector<int64_t> lanes, yd;
  std::vector<double> cost;
  for (auto& c : p->classes) {
    c.generic = c.native_generic;
    c.cls = c.spec_cls;
    const bool few = c.count * 4 <= (int64_t)ctx->ncu * 256;
    // few deep quartets of a class whose register kernel was measured slower
    // there than the generic kernel (with its quartets split over blocks),
    // in a plan that is mostly this class: the split generic passes fill
    // the chip's LDS, and beside other classes (a mixed batch) the wave per
    // quartet, which leaves most of the chip to them, measured faster overall
    if (p->wave_split && p->few_generic && !c.generic && !c.fp32 && few && c.nprimq >= 16 * c.count &&
        c.gidx >= 0 && l2g_gen_classes[c.spec_cls].few_generic && c.count * 2 >= p->nq) {
      c.generic = true;
      c.cls = c.gidx;
    }
    c.coop = p->wave_split && !c.generic && few && c.nprimq >= 16 * c.count;
    if (!c.generic && !c.coop && !c.fp32 && c.gidx >= 0 && few && c.nprimq >= L2G_GENERIC_MIN_DEPTH * c.count &&
        class_l(c) >= L2G_GENERIC_MIN_L) {
      c.generic = true;
      c.cls = c.gidx;
    }
    // the generic kernel runs a block per quartet, no lane scratch, a few
    // blocks per CU (LDS bound): up to 4 quartets per CU leave the chip
    // underfilled, so such classes count as small and run beside each other
    // on the side streams
    lanes.push_back(c.generic ? c.count * 64 : (c.coop ? 64 * c.count : c.count));
    yd.push_back(c.generic ? 0 : (c.fp32 ? l2g_gen_classes32 : l2g_gen_classes)[c.cls].ybuf_doubles);
  }
  plan_pieces(p, ctx);
  for (size_t i = 0; i < p->classes.size(); ++i)
    if (p->classes[i].npieces) lanes[i] = p->classes[i].npieces * 64;
  // expected time for the side-stream packing: recursion values per
  // primitive quartet times primitive quartets, at each kernel kind's rate
  // (ns per value, measured on an s-f batch: the generic kernel 0.005 with
  // its quartets split over blocks, 0.03 without; a wave per quartet 0.0025;
  // a lane per quartet 0.02), plus 100 ns per quartet
  for (const auto& c : p->classes) {
    const double vals = c.generic ? (double)l2g_gen_generic[c.cls].nvals
                                  : (double)(c.fp32 ? l2g_gen_classes32 : l2g_gen_classes)[c.cls].nodes;
    const double rate = c.generic ? (c.npieces ? 0.005 : 0.03) : (c.coop ? 0.0025 : 0.02);
    cost.push_back((double)c.nprimq * vals * rate + 100.0 * (double)c.count);
  }
  std::vector<ClassSlot> slots;
  // side regions up to 4 GiB of scratch beyond the shared one
  const int64_t total = plan_slots(slots, lanes, yd, ctx->ncu, p->block, l2g_context_s::kSide,
                                   p->ymax * (int64_t)ctx->ncu * 8 * p->block + (int64_t(1) << 29), &cost);
  for (size_t i = 0; i < p->classes.size(); ++i) p->classes[i].slot = slots[i];
  p->ybuf.alloc(std::max<int64_t>(total, 1) * sizeof(double));
}

}  // namespace

extern "C" {

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
    p->nz.alloc(std::max<int64_t>(p->nq, 1));
    A.nz = p->nz.as<signed char>();
    if (!p->t0) {
      HIPCHK(hipEventCreate(&p->t0));
      HIPCHK(hipEventCreate(&p->t1));
    }
    // a replay reuses the captured kernel arguments: recapture for another
    // output buffer or after the solid-harmonic ordering changed
    const bool capture =
        p->use_graph && !(p->gexec && p->graph_out == out && p->graph_gauss == (int)ctx->shg_gaussian);
    if (p->use_graph && !capture) {  // replay
      HIPCHK(hipEventRecord(p->t0, ctx->stream));
      HIPCHK(hipGraphLaunch(p->gexec, ctx->stream));
      HIPCHK(hipEventRecord(p->t1, ctx->stream));
      return;
    }
    if (capture) {
      p->drop_graph();
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    } else {
      HIPCHK(hipEventRecord(p->t0, ctx->stream));
    }
    // classes too small to fill the chip run beside each other on the side
    // streams (not under capture, where the launches stay in order)
    bool sides = false;
    for (auto& c : p->classes) sides |= c.slot.side >= 0;
    sides &= !capture;
    // attenuated passes run on the class's erf-attenuated register kernel
    // (a separate instantiation, so the Coulomb kernels' primitive loop
    // carries no operator branch) where one is built, else on the generic
    // kernel, whose core-integral step takes omega; the Coulomb pass of erfc
    // runs on the Coulomb kernels.
    // Passes (attenuation, subtract): Coulomb; erf = one attenuated pass
    // (omega = 0 gives exact zeros as Coulomb minus Coulomb); erfc =
    // Coulomb minus the erf pass, both per class on the class's stream
    std::vector<std::pair<double, int>> passes;
    if (p->oper == 0) passes = {{0.0, 0}};
    else if (p->oper == 1) passes = p->omega > 0 ? std::vector<std::pair<double, int>>{{p->omega, 0}}
                                                 : std::vector<std::pair<double, int>>{{0.0, 0}, {0.0, 1}};
    else passes = p->omega > 0 ? std::vector<std::pair<double, int>>{{0.0, 0}, {p->omega, 1}}
                               : std::vector<std::pair<double, int>>{{0.0, 0}};
    auto launch_classes = [&] {
    for (auto& c : p->classes) {
    for (size_t ps = 0; ps < passes.size(); ++ps) {
      A.omega = passes[ps].first;
      A.store_sub = passes[ps].second;
      hipStream_t cs = sides && c.slot.side >= 0 ? ctx->side[c.slot.side] : ctx->stream;
      A.ybuf = p->ybuf.as<double>() + c.slot.yoff;
      A.ystride = c.slot.ystride;
      A.bra_pair = p->bra.as<int>() + c.first;
      A.ket_pair = p->ket.as<int>() + c.first;
      A.meta = p->meta.as<int>() + c.first;
      A.out_off = p->off.as<long long>() + c.first;
      A.qidx = p->qidx.as<long long>() + c.first;
      A.n = (int)c.count;
      if (!capture && ps == 0) HIPCHK(hipEventRecord(c.ev0, cs));
      int erf_idx = -1;  // the class's erf-attenuated register kernel, for an attenuated pass
      if (A.omega > 0 && !c.generic && !c.fp32) {
        const L2gClassInfo& ci = l2g_gen_classes[c.cls];
        erf_idx = ctx->classerf_of[ci.la][ci.lb][ci.lc][ci.ld];
      }
      if (erf_idx >= 0) {
        const L2gClassInfo& ci = l2g_gen_classes_erf[erf_idx];
        A.coop = c.coop ? 1 : 0;
        ci.launch(A, c.slot.grid, p->block, cs);
        LAUNCHCHK(std::string("erf ERI kernel of class ") + std::to_string(c.cls));
      } else if (c.generic || A.omega > 0) {  // the attenuated pass of erf / erfc
        // one block per quartet, persistent over the class
        if (c.gidx < 0) throw L2gError(L2G_E_LMAX, "no generic-kernel program for this class");
        const L2gGenericClass& gc = l2g_gen_generic[c.gidx];
        if (c.generic && c.npieces) {  // pieces of the primitive loops, then the sum per quartet
          L2gEriArgs P = A;
          P.bra_pair = p->gp_bra.as<int>() + c.piece_off;
          P.ket_pair = p->gp_ket.as<int>() + c.piece_off;
          P.meta = p->gp_meta.as<int>() + c.piece_off;
          P.qidx = p->gp_qidx.as<long long>() + c.piece_off;
          P.q_prange = p->gp_range.as<int2>() + c.piece_off;
          P.n = (int)c.npieces;
          P.gpart_mode = 1;
          P.gpart = p->gpart.as<double>() + c.gpart_off;
          P.gpart_stride = c.gstride;
          launch_generic(kGenericEri, P, gc, ctx->gprog.as<int4>(), ctx->gidx.as<int>(),
                         (int)std::min<int64_t>((int64_t)ctx->ncu * 8, c.npieces), cs);
          P = A;
          P.gpart_mode = 2;
          P.gpart = p->gpart.as<double>() + c.gpart_off;
          P.gpart_stride = c.gstride;
          P.gpart_first = p->gp_first.as<int>() + c.pfirst_off;
          launch_generic(kGenericEri, P, gc, ctx->gprog.as<int4>(), ctx->gidx.as<int>(),
                         (int)std::min<int64_t>((int64_t)ctx->ncu * 8, c.count), cs);
        } else {
          const int grid = (int)std::min<int64_t>((int64_t)ctx->ncu * 8, c.count);
          launch_generic(kGenericEri, A, gc, ctx->gprog.as<int4>(), ctx->gidx.as<int>(), grid, cs);
        }
      } else {
        const L2gClassInfo& ci = (c.fp32 ? l2g_gen_classes32 : l2g_gen_classes)[c.cls];
        A.coop = c.coop ? 1 : 0;
        // exactly the slot's grid: the scratch region holds grid * block lanes
        ci.launch(A, c.slot.grid, p->block, cs);
        LAUNCHCHK(std::string("ERI kernel of class ") + std::to_string(c.cls));
      }
      if (!capture && ps + 1 == passes.size()) HIPCHK(hipEventRecord(c.ev1, cs));
    }
    }
    };
    // the side streams are joined back also when a launch fails
    auto launch_all = [&] {
      if (sides) fork_sides(*ctx);
      try {
        launch_classes();
      } catch (...) {
        if (sides) {
          try {
            join_sides(*ctx);
          } catch (...) {
          }
        }
        throw;
      }
      if (sides) join_sides(*ctx);
    };
    if (!capture) {
      launch_all();
      HIPCHK(hipEventRecord(p->t1, ctx->stream));
      return;
    }
    try {
      launch_all();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(ctx->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(ctx->stream, &p->graph));
    HIPCHK(hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
    p->graph_out = out;
    p->graph_gauss = (int)ctx->shg_gaussian;
    HIPCHK(hipEventRecord(p->t0, ctx->stream));
    HIPCHK(hipGraphLaunch(p->gexec, ctx->stream));
    HIPCHK(hipEventRecord(p->t1, ctx->stream));
  });
}

int l2g_eri_plan_set_graph(l2g_eri_plan p, int on) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null plan");
    (void)hipSetDevice(p->basis->ctx->device);
    HIPCHK(hipStreamSynchronize(p->basis->ctx->stream));
    p->use_graph = on != 0;
    if (!p->use_graph) p->drop_graph();
  });
}

int l2g_eri_plan_set_operator(l2g_eri_plan p, int oper, double omega) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null plan");
    if (oper < 0 || oper > 2 || !(omega >= 0)) throw std::invalid_argument("l2g_eri_plan_set_operator: bad operator or omega");
    l2g_context_s* ctx = p->basis->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    p->oper = oper;
    p->omega = omega;
    p->drop_graph();
  });
}

int l2g_eri_plan_set_wave_split(l2g_eri_plan p, int on) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null plan");
    l2g_context_s* ctx = p->basis->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (on < 0 || on > 2) throw std::invalid_argument("wave split mode must be 0, 1 or 2");
    p->wave_split = on != 0;
    p->few_generic = on != 2;
    p->drop_graph();
    plan_layout(p, ctx);
  });
}

int l2g_eri_plan_elapsed(l2g_eri_plan p, float* ms) {
  return guarded([&] {
    if (!p || !ms) throw std::invalid_argument("null argument");
    if (!p->t0) throw L2gError(L2G_E_STATE, "plan not executed");
    HIPCHK(hipEventSynchronize(p->t1));
    HIPCHK(hipEventElapsedTime(ms, p->t0, p->t1));
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
      int la, lb, lc, ld;
      if (c.generic) {
        const auto& g = l2g_gen_generic[c.cls];
        la = g.la, lb = g.lb, lc = g.lc, ld = g.ld;
      } else {
        const auto& ci = (c.fp32 ? l2g_gen_classes32 : l2g_gen_classes)[c.cls];
        la = ci.la, lb = ci.lb, lc = ci.lc, ld = ci.ld;
      }
      if (code)
        code[i] = la * 1000 + lb * 100 + lc * 10 + ld + (c.fp32 ? 10000 : 0) + (c.generic ? 20000 : 0) +
                  (c.coop ? 100000 : 0);
      if (nquart) nquart[i] = c.count;
      if (nprimq) nprimq[i] = c.nprimq;
      if (ms) {
        ms[i] = p->use_graph ? -1.0f : event_ms_or(c.ev0, c.ev1, -1.0f);
      }
    }
  });
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

int l2g_eri_plan_destroy(l2g_eri_plan p) {
  if (p && !g_exiting) {
    l2g_basis_s* B = p->basis;
    l2g_context_s* ctx = B->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete p;
    --ctx->live;
    basis_release(B);  // the plan's hold on its basis
  }
  return L2G_OK;
}

}  // extern "C"

namespace l2g_impl {

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
  // buckets: generated Schwarz kernels, then generic classes (ab|ab)
  const int nsch = l2g_gen_nschwarz;
  std::vector<std::vector<int>> pid(nsch + l2g_gen_ngeneric), meta(nsch + l2g_gen_ngeneric);
  std::vector<std::vector<long long>> idx(nsch + l2g_gen_ngeneric);
  for (int64_t q = 0; q < npairs; ++q) {
    int a = pairs[2 * q], b = pairs[2 * q + 1];
    if (a < 0 || b < 0 || a >= B->nshell || b >= B->nshell) throw std::invalid_argument("shell index out of range");
    if (B->swap_pair(a, b)) std::swap(a, b);
    int c = ctx->force_generic ? -1 : ctx->schwarz_of[B->l[a]][B->l[b]];
    if (c < 0) {
      const int g = ctx->generic_of[B->l[a]][B->l[b]][B->l[a]][B->l[b]];
      if (g < 0) throw L2gError(L2G_E_LMAX, "no Schwarz kernel for this shell pair");
      c = nsch + g;
    }
    const int pa = (B->pure[a] && B->l[a] > 0) ? 1 : 0, pbit = (B->pure[b] && B->l[b] > 0) ? 1 : 0;
    pid[c].push_back(pb.get(a, b));
    meta[c].push_back((pa << 3) | (pbit << 4) | (pa << 5) | (pbit << 6));
    idx[c].push_back(q);
  }
  std::vector<int> allp, allm;
  std::vector<long long> alli;
  std::vector<std::pair<int, int64_t>> ranges;
  int64_t ymax = 0, maxn = 0;
  for (int c = 0; c < (int)pid.size(); ++c) {
    if (pid[c].empty()) continue;
    ranges.push_back({c, (int64_t)allp.size()});
    allp.insert(allp.end(), pid[c].begin(), pid[c].end());
    allm.insert(allm.end(), meta[c].begin(), meta[c].end());
    alli.insert(alli.end(), idx[c].begin(), idx[c].end());
    if (c < nsch) ymax = std::max<int64_t>(ymax, l2g_gen_schwarz[c].ybuf_doubles);
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
  // ybuf columns for the largest launch: one lane per pair, or a wave per pair
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((64 * maxn + block - 1) / block, (int64_t)ctx->ncu * 8));
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
    if (c < nsch) {
      // deep pairs: one wave per pair splits the (p|p) primitive loop, so a
      // contracted pair's latency is not its whole serial loop
      double mean = 0;
      for (int64_t i = 0; i < n; ++i) {
        const double q = (double)pb.np[pid[c][i]];
        mean += q * q;
      }
      mean /= (double)n;
      A.coop = mean >= 16.0 ? 1 : 0;
      const int64_t want = A.coop ? (64 * n + block - 1) / block : (n + block - 1) / block;
      l2g_gen_schwarz[c].launch(A, (int)std::min<int64_t>((int64_t)ctx->ncu * 8, want), block, st);
      LAUNCHCHK(std::string("Schwarz kernel of class ") + std::to_string(c));
    } else {
      launch_generic(kGenericSchwarz, A, l2g_gen_generic[c - nsch], ctx->gprog.as<int4>(), ctx->gidx.as<int>(),
                     (int)std::min<int64_t>((int64_t)ctx->ncu * 8, n), st);
    }
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

}  // namespace l2g_impl

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


// ===========================================================================
// First geometric derivatives (deriv_order = 1)
// ===========================================================================
// Engine(Operator::coulomb, max_nprim, max_l, 1)::compute returns, per
// quartet, 12 shell sets d/dA_x .. d/dD_z for the shells' centres in caller
// order (CartesianDerivIterator<4>, include/libint2/util/intpart_iter.h:
// 89-103, 161-182).  The plan evaluates them from ERIs of an augmented
// basis, shells 0..n-1 the caller's shells made Cartesian, n..2n-1 raised
// (l+1, coefficients 2 alpha c), 2n..3n-1 lowered (l-1; an unused s copy for
// l = 0), in one ERI plan over the 8 nq quartets with one shell raised or
// lowered; a combination kernel forms the derivatives and applies the
// solid-harmonic transform of the pure centres.
struct l2g_deriv1_plan_s {
  l2g_context_s* ctx = nullptr;
  l2g_basis aug = nullptr;
  l2g_eri_plan inner = nullptr;
  int64_t nq = 0, total = 0;
  std::vector<int64_t> offsets;
  std::vector<int> lq;  // l of the four shells of each quartet, 8 bits each
  DevBuf raw_off, out_off, qinfo, out;
  ~l2g_deriv1_plan_s() {
    if (inner) l2g_eri_plan_destroy(inner);
    if (aug) l2g_basis_destroy(aug);
  }
};

namespace {

L2gDerivTab deriv_tab(bool gaussian) {
  L2gDerivTab t;
  std::memset(&t, 0, sizeof(t));
  for (int l = 0; l <= 3; ++l)
    for (int f = 0; f < 2 * l + 1; ++f) {
      // output position f -> m (Standard: m = f - l; Gaussian: 0, +1, -1, +2, ...)
      const int m = gaussian ? (f == 0 ? 0 : ((f & 1) ? (f + 1) / 2 : -f / 2)) : f - l;
      int c = 0, k = 0;
      for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y, ++c) {
          const double v = l2g_solid_harmonics_coeff(l, m, x, y, l - x - y);
          if (v == 0.0) continue;
          if (k == 4) throw std::logic_error("solid-harmonic row with more than 4 terms");
          t.cidx[l][f][k] = (signed char)c;
          t.coef[l][f][k] = v;
          ++k;
        }
      t.nnz[l][f] = (signed char)k;
    }
  return t;
}

}  // namespace

extern "C" {

int l2g_eri_deriv1_plan_create(l2g_basis B, int64_t nq, const int32_t* quartets, double precision, int screening,
                               l2g_deriv1_plan* out) {
  return guarded([&] {
    if (!B) throw L2gError(L2G_E_STATE, "null basis");
    if (!out || nq < 0 || (nq > 0 && !quartets)) throw std::invalid_argument("l2g_eri_deriv1_plan_create: bad arguments");
    l2g_context_s* ctx = B->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = B->nshell;
    for (int64_t i = 0; i < 4 * nq; ++i)
      if (quartets[i] < 0 || quartets[i] >= n) throw std::invalid_argument("shell index out of range");
    // raised shells need l + 1 <= the kernels' limit; the combination kernel
    // stages at most 6^4 Cartesian elements (l <= 2)
    const int dlmax = std::min(lmax_supported() - 1, 2);
    for (int64_t i = 0; i < 4 * nq; ++i)
      if (B->l[quartets[i]] > dlmax)
        throw L2gError(L2G_E_LMAX, "first derivatives are provided up to l = " + std::to_string(dlmax) +
                                       " (angular momentum " + std::to_string(B->l[quartets[i]]) + ")");
    auto p = std::make_unique<l2g_deriv1_plan_s>();
    p->ctx = ctx;
    p->nq = nq;
    // augmented basis; shells above the derivative limit are never referenced
    // by the (validated) quartets, and their raised copies keep their own l
    std::vector<int> al, ap, an;
    std::vector<double> aa, ac, ao;
    for (int kind = 0; kind < 3; ++kind)
      for (int s = 0; s < n; ++s) {
        const int l = B->l[s];
        al.push_back(kind == 1 ? (l <= dlmax ? l + 1 : l) : (kind == 2 ? std::max(l - 1, 0) : l));
        ap.push_back(0);
        an.push_back(B->nprim[s]);
        for (int k = B->off[s]; k < B->off[s + 1]; ++k) {
          aa.push_back(B->alpha[k]);
          ac.push_back(kind == 1 ? 2.0 * B->alpha[k] * B->coeff[k] : B->coeff[k]);
        }
        for (int x = 0; x < 3; ++x) ao.push_back(B->origin[3 * s + x]);
      }
    int rc = l2g_basis_create(ctx, 3 * n, al.data(), ap.data(), an.data(), aa.data(), ac.data(), ao.data(), &p->aug);
    if (rc) throw L2gError(rc, g_err);
    std::vector<int32_t> big((size_t)32 * nq);
    for (int k = 0; k < 8; ++k)
      for (int64_t q = 0; q < nq; ++q)
        for (int j = 0; j < 4; ++j)
          big[4 * ((size_t)k * nq + q) + j] = quartets[4 * q + j] + (j == (k & 3) ? (k < 4 ? n : 2 * n) : 0);
    rc = l2g_eri_plan_create(p->aug, 8 * nq, big.data(), precision, screening, 1.0, 0, &p->inner);
    if (rc) throw L2gError(rc, g_err);
    p->offsets.resize(nq);
    p->lq.resize(nq);
    std::vector<int> qinfo(nq);
    int64_t total = 0;
    for (int64_t q = 0; q < nq; ++q) {
      const int32_t* s = quartets + 4 * q;
      int info = 0, lw = 0;
      int64_t sz = 1;
      for (int k = 0; k < 4; ++k) {
        const int l = B->l[s[k]];
        const bool pure = B->pure[s[k]] && l > 0;
        info |= l << (4 * k);
        if (pure) info |= 1 << (16 + k);
        lw |= l << (8 * k);
        sz *= nfun(l, pure ? 1 : 0);
      }
      qinfo[q] = info;
      p->lq[q] = lw;
      p->offsets[q] = total;
      total += sz;
    }
    p->total = total;
    std::vector<long long> ro(p->inner->offsets.begin(), p->inner->offsets.end());
    std::vector<long long> oo(p->offsets.begin(), p->offsets.end());
    upload(p->raw_off, ro, ctx->stream);
    upload(p->out_off, oo, ctx->stream);
    upload(p->qinfo, qinfo, ctx->stream);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ++ctx->live;
    *out = p.release();
  });
}

int l2g_eri_deriv1_plan_size(l2g_deriv1_plan p, int64_t* total) {
  return guarded([&] {
    if (!p || !total) throw std::invalid_argument("null argument");
    *total = p->total;
  });
}

int l2g_eri_deriv1_plan_offsets(l2g_deriv1_plan p, int64_t* offsets) {
  return guarded([&] {
    if (!p || !offsets) throw std::invalid_argument("null argument");
    std::memcpy(offsets, p->offsets.data(), p->offsets.size() * sizeof(int64_t));
  });
}

int l2g_eri_deriv1_plan_execute(l2g_deriv1_plan p, double* d_out) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null plan");
    l2g_context_s* ctx = p->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    double* out = d_out;
    if (!out) {
      p->out.alloc(std::max<int64_t>(12 * p->total, 1) * sizeof(double));
      out = p->out.as<double>();
    }
    int rc = l2g_eri_plan_execute(p->inner, nullptr);
    if (rc) throw L2gError(rc, g_err);
    if (p->nq == 0) return;
    launch_deriv1_combine(p->inner->out.as<double>(), p->raw_off.as<long long>(), p->out_off.as<long long>(),
                          p->qinfo.as<int>(), p->nq, p->total, out, deriv_tab(ctx->shg_gaussian != 0), ctx->stream);
  });
}

int l2g_eri_deriv1_plan_download(l2g_deriv1_plan p, double* host_out, int8_t* nonzero) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null plan");
    l2g_context_s* ctx = p->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) throw L2gError(L2G_E_LAUNCH, std::string("kernel execution failed: ") + hipGetErrorString(e));
    if (host_out && p->total > 0) {
      if (!p->out.p) throw L2gError(L2G_E_STATE, "plan has not been executed into its own buffer");
      HIPCHK(hipMemcpy(host_out, p->out.p, 12 * p->total * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (nonzero && p->nq > 0) {
      std::vector<int8_t> nz(8 * p->nq);
      int rc = l2g_eri_plan_download(p->inner, nullptr, nz.data());
      if (rc) throw L2gError(rc, g_err);
      for (int64_t q = 0; q < p->nq; ++q) {
        int8_t v = 0;
        for (int k = 0; k < 4; ++k) {
          v |= nz[(size_t)k * p->nq + q];
          if ((p->lq[q] >> (8 * k)) & 255) v |= nz[(size_t)(4 + k) * p->nq + q];
        }
        nonzero[q] = v ? 1 : 0;
      }
    }
  });
}

int l2g_eri_deriv1_plan_destroy(l2g_deriv1_plan p) {
  if (p && !g_exiting) {
    l2g_context_s* ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete p;
    --ctx->live;
  }
  return L2G_OK;
}

}  // extern "C"

// ===========================================================================
// One-body integrals
// ===========================================================================
extern "C" {

int l2g_onebody(l2g_basis B, int ncharge, const double* charges, const double* xyz, double* S, double* T, double* V) {
  return guarded([&] {
    if (!B) throw L2gError(L2G_E_STATE, "null basis");
    if (ncharge < 0 || (ncharge > 0 && (!charges || !xyz))) throw std::invalid_argument("l2g_onebody: bad arguments");
    if (V && ncharge == 0) throw std::invalid_argument("l2g_onebody: nuclear attraction needs at least one charge");
    l2g_context_s* ctx = B->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = B->nshell;
    std::vector<long long> bf0(n + 1, 0);
    for (int s = 0; s < n; ++s) {
      if (B->l[s] > 3) throw L2gError(L2G_E_LMAX, "one-body integrals are provided up to l = 3");
      bf0[s + 1] = bf0[s] + nfun(B->l[s], (B->pure[s] && B->l[s] > 0) ? 1 : 0);
    }
    const long long nbf = bf0[n];
    if (n == 0) return;
    hipStream_t st = ctx->stream;
    DevBuf dl, dp, doff, da, dc, dO, dbf, dq, dqx, dS, dT, dV;
    upload(dl, B->l, st);
    upload(dp, B->pure, st);
    upload(doff, B->off, st);
    upload(da, B->alpha, st);
    upload(dc, B->coeff, st);
    upload(dO, B->origin, st);
    upload(dbf, bf0, st);
    std::vector<double> q(charges, charges + ncharge), qx(xyz, xyz + 3 * ncharge);
    upload(dq, q, st);
    upload(dqx, qx, st);
    const size_t bytes = (size_t)nbf * nbf * sizeof(double);
    L2gOneBodyArgs A;
    std::memset(&A, 0, sizeof(A));
    A.l = dl.as<int>();
    A.pure = dp.as<int>();
    A.poff = doff.as<int>();
    A.alpha = da.as<double>();
    A.coeff = dc.as<double>();
    A.origin = dO.as<double>();
    A.bf0 = dbf.as<long long>();
    A.nshell = n;
    A.nbf = nbf;
    A.ncharge = ncharge;
    A.q = dq.as<double>();
    A.qxyz = dqx.as<double>();
    A.boys = ctx->boys.as<double>();
    A.gauss = ctx->shg_gaussian;
    if (S) dS.alloc(bytes), A.S = dS.as<double>();
    if (T) dT.alloc(bytes), A.T = dT.as<double>();
    if (V) dV.alloc(bytes), A.V = dV.as<double>();
    launch_onebody(A, st);
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) throw L2gError(L2G_E_LAUNCH, std::string("kernel execution failed: ") + hipGetErrorString(e));
    if (S) HIPCHK(hipMemcpy(S, dS.p, bytes, hipMemcpyDeviceToHost));
    if (T) HIPCHK(hipMemcpy(T, dT.p, bytes, hipMemcpyDeviceToHost));
    if (V) HIPCHK(hipMemcpy(V, dV.p, bytes, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
