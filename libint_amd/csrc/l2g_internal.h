// Internals shared by the host translation units of the C ABI (l2g_host.cpp,
// l2g_fock.cpp): errors, device buffers, context, basis and shell-pair data.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
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
// erf-attenuated ERI variants (same plans and scratch layout as l2g_gen_classes)
extern const L2gClassInfo l2g_gen_classes_erf[];
extern const int l2g_gen_nclasses_erf;
extern const int l2g_gen_lmax;
// programs of the generic kernel (codegen/gen_generic.py)
extern const L2gGenericClass l2g_gen_generic[];
extern const int l2g_gen_ngeneric;
extern const int l2g_gen_generic_lmax;
extern const int l2g_gen_generic_prog[];
extern const int l2g_gen_generic_nprog;
extern const int l2g_gen_generic_idx[];
extern const int l2g_gen_generic_nidx;


namespace l2g_impl {


inline thread_local std::string g_err;

// Process teardown.  The first l2g_init registers an atexit handler; atexit
// handlers run in reverse order of registration, so it runs before the HIP
// runtime (loaded, and registered, earlier) tears itself down.  The handler
// stops the threads of every live request server and sets g_exiting; from
// then on the destroy entry points and the device-buffer destructors touch
// nothing on the device: the process is ending and its device resources go
// with it.
inline std::atomic<bool> g_exiting{false};
void register_exit_handler();  // l2g_host.cpp
void stop_servers_at_exit();   // l2g_fock.cpp

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

// error check after a kernel launch, naming the kernel
#define LAUNCHCHK(name)                                                                             \
  do {                                                                                              \
    hipError_t e_ = hipGetLastError();                                                              \
    if (e_ != hipSuccess)                                                                           \
      throw L2gError(e_ == hipErrorOutOfMemory ? L2G_E_OOM : L2G_E_LAUNCH,                        \
                     std::string("launch of ") + (name) + ": " + hipGetErrorString(e_));           \
  } while (0)

template <class F>
inline int guarded(F&& f) {
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

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
inline int nfun(int l, int pure) { return pure ? 2 * l + 1 : ncart(l); }


// ---------------------------------------------------------------------------
// normalization and solid harmonics
// ---------------------------------------------------------------------------
inline double double_factorial_km1(int k) {  // (k-1)!!
  double r = 1;
  for (int i = k - 1; i > 1; i -= 2) r *= i;
  return r;
}
inline double factorial(int n) {
  double r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
inline double choose(int n, int k) {
  if (k < 0 || k > n) return 0;
  return factorial(n) / (factorial(k) * factorial(n - k));
}
inline int sgn_parity(int i) { return (i & 1) ? -1 : 1; }

// ---------------------------------------------------------------------------
// objects behind the opaque handles
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  // owning: moved (std::swap included), never copied
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p, bytes = o.bytes;
      o.p = nullptr, o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() {
    if (p && !g_exiting) (void)hipFree(p);
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

// Elapsed time between two events where it may legitimately be unavailable
// (events not recorded yet): dflt then, and the runtime's last-error state
// is cleared so that a later launch check does not report the failure.
inline float event_ms_or(hipEvent_t a, hipEvent_t b, float dflt) {
  float t = 0;
  if (hipEventElapsedTime(&t, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return dflt;
  }
  return t;
}

template <class T>
inline void upload(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  d.alloc(std::max<size_t>(h.size() * sizeof(T), 16));
  if (!h.empty()) HIPCHK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
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

// Launch slots of a batch of classes.  A class whose lanes fit one wave per
// SIMD of the chip cannot fill it alone: such classes run beside each other
// on the side streams, each with its own lane-private scratch region; the
// larger classes run in order on the main stream and share one region.
// lanes[i] is the class's lanes (0: class absent), ydoubles[i] its scratch
// doubles per lane; a class launches exactly slots[i].grid blocks of `block`
// threads (with ystride = grid * block).  Returns the scratch doubles needed;
// when the side regions would exceed cap_doubles every class runs on the
// main stream (side = -1) in the shared region.  cost (optional): each
// class's expected time; the small classes then go to the side streams
// longest first, each to the stream with the least expected time so far (so
// a long class starts at once and the others pack beside it), else in turn.
struct ClassSlot {
  int64_t yoff = 0, ystride = 0;
  int grid = 1, side = -1;
};
inline int64_t plan_slots(std::vector<ClassSlot>& slots, const std::vector<int64_t>& lanes,
                          const std::vector<int64_t>& ydoubles, int ncu, int block, int nside, int64_t cap_doubles,
                          const std::vector<double>* cost = nullptr) {
  const size_t n = lanes.size();
  slots.assign(n, ClassSlot());
  const int64_t maxblocks = (int64_t)ncu * 8, layer = (int64_t)ncu * 256;
  int64_t shared = 0, sides = 0;
  int nsmall = 0;
  for (size_t i = 0; i < n; ++i) {
    if (lanes[i] <= 0) continue;
    ClassSlot& s = slots[i];
    s.grid = (int)std::max<int64_t>(1, std::min<int64_t>(maxblocks, (lanes[i] + block - 1) / block));
    s.ystride = (int64_t)s.grid * block;
    if (lanes[i] <= layer && nside > 0) {
      s.side = nsmall++ % nside;  // reassigned below when costs are given
      s.yoff = -1;  // placed below
      sides += ydoubles[i] * s.ystride;
    } else {
      shared = std::max(shared, ydoubles[i] * s.ystride);
    }
  }
  if (nsmall && shared + sides > cap_doubles) {  // sequential: one shared region
    for (size_t i = 0; i < n; ++i) {
      if (lanes[i] <= 0) continue;
      slots[i].side = -1, slots[i].yoff = 0;
      shared = std::max(shared, ydoubles[i] * slots[i].ystride);
    }
    return shared;
  }
  if (cost && nsmall) {
    std::vector<size_t> order;
    for (size_t i = 0; i < n; ++i)
      if (lanes[i] > 0 && slots[i].side >= 0) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return (*cost)[a] > (*cost)[b]; });
    std::vector<double> load(nside, 0.0);
    for (size_t i : order) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      slots[i].side = k;
      load[k] += (*cost)[i];
    }
  }
  int64_t off = shared;
  for (size_t i = 0; i < n; ++i)
    if (lanes[i] > 0 && slots[i].side >= 0) {
      slots[i].yoff = off;
      off += ydoubles[i] * slots[i].ystride;
    }
  return off;
}

// Forks the side streams from the main stream (they start after the work
// queued on it so far) and, at the end, joins them back.
inline void fork_sides(const l2g_context_s& ctx);
inline void join_sides(const l2g_context_s& ctx);

}  // namespace l2g_impl

struct l2g_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // side streams for classes that run beside each other (l2g_impl::plan_slots),
  // forked from and joined back into `stream` with events
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr, join[kSide] = {};
  l2g_impl::DevBuf boys;
  int shg_gaussian = 0;
  int ncu = 256;
  int class_of[8][8][8][8];  // index into l2g_gen_classes
  int fock_of[8][8][8][8];   // index into l2g_gen_fock (its own order)
  int class32_of[8][8][8][8];  // index into l2g_gen_classes32
  int classerf_of[8][8][8][8]; // index into l2g_gen_classes_erf
  int fock32_of[8][8][8][8];   // index into l2g_gen_fock32
  int schwarz_of[8][8];
  int generic_of[8][8][8][8];  // index into l2g_gen_generic
  bool force_generic = false;  // every class through the generic kernel (testing)
  // FP32 routing of mixed batches: 0, only classes whose FP32 variant was
  // measured faster than their FP64 kernel (L2gClassInfo::fp32_pays); 1,
  // every class with an FP32 variant
  int fp32_policy = 0;
  l2g_impl::DevBuf gprog, gidx;  // device copies of the generic programs
  // bases, plans and Fock builders created on this context and not yet
  // destroyed; l2g_finalize refuses while any is alive
  std::atomic<int> live{0};
};
namespace l2g_impl {
inline void fork_sides(const l2g_context_s& ctx) {
  HIPCHK(hipEventRecord(ctx.fork, ctx.stream));
  for (int k = 0; k < l2g_context_s::kSide; ++k) HIPCHK(hipStreamWaitEvent(ctx.side[k], ctx.fork, 0));
}
inline void join_sides(const l2g_context_s& ctx) {
  for (int k = 0; k < l2g_context_s::kSide; ++k) {
    HIPCHK(hipEventRecord(ctx.join[k], ctx.side[k]));
    HIPCHK(hipStreamWaitEvent(ctx.stream, ctx.join[k], 0));
  }
}
}  // namespace l2g_impl


struct l2g_basis_s {
  l2g_context_s* ctx;
  // owners: the caller's handle plus every plan and Fock builder made on the
  // basis; the basis is freed when the last of them lets go, so a plan may
  // outlive l2g_basis_destroy of its basis (as happens when a garbage
  // collector finalizes both in either order)
  std::atomic<int> refs{1};
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
  // Shells below f keep the caller's order, as the reference: when both
  // shells of a pair mix tight and diffuse primitives neither centre is
  // reliably better, and on d pairs the caller's order measured closer to
  // the exact integrals; on f pairs the tighter shell cut the Fock error of
  // an f-recipe build against the exact one from 2e-6 to 3e-11.
  bool second_is_centre(int i, int j) {
    if (i == j || l[i] != l[j] || l[i] < 3) return false;
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

namespace l2g_impl {
inline void basis_retain(l2g_basis_s* b) { ++b->refs; }
inline void basis_release(l2g_basis_s* b) {
  if (b && --b->refs == 0) delete b;
}
}  // namespace l2g_impl

// FP32 selection of a mixed batch: a quartet runs in FP32 when its Schwarz
// bound K_ab K_cd is below the caller's threshold and its pre-HRR bound
// S_ab S_cd (schwarz_impl) is below this factor times the threshold, which
// keeps eps_fp32 S_ab S_cd, the scale of the FP32 rounding that reaches an
// integral, near eps_fp32 times the threshold.
#ifndef L2G_FP32_HRR_FACTOR
#define L2G_FP32_HRR_FACTOR 16.0
#endif


namespace l2g_impl {

void schwarz_impl(l2g_basis B, int64_t npairs, const int32_t* pairs, double precision, int norm, double* K,
                  double* S);

// l2g_aux.hip: in-place Fock symmetrisation of one N x N block (J into jg,
// K into kg, G = 2J - K into g)
// hoff: 0 for FP64 accumulators; else the fixed-point accumulators' offset
// from a low word to its high word (l2g_fx_add)
void launch_fock_finish(double* jg, double* kg, double* g, long long N, long long hoff, hipStream_t s);

// Solid-harmonic rows in output order (Standard or Gaussian) for l <= 3: the
// Cartesian components and coefficients of output function f of a pure
// shell of angular momentum l are cidx/coef[l][f][0 .. nnz[l][f]).
struct L2gDerivTab {
  signed char nnz[4][7];
  signed char cidx[4][7][4];
  double coef[4][7][4];
};

// l2g_generic.hip: one-body integrals over every shell pair of a basis.
// Per-shell arrays (poff = first primitive, size nshell + 1; bf0 = first
// function), point charges q at qxyz; S, T, V (any may be null) are nbf x
// nbf row-major on the device.
struct L2gOneBodyArgs {
  const int* l;
  const int* pure;
  const int* poff;
  const double* alpha;
  const double* coeff;
  const double* origin;
  const long long* bf0;
  int nshell;
  long long nbf;
  int ncharge;
  const double* q;
  const double* qxyz;
  const double* boys;
  int gauss;
  double* S;
  double* T;
  double* V;
};
void launch_onebody(const L2gOneBodyArgs& A, hipStream_t s);

// l2g_aux.hip: first-derivative shell sets from the raised and lowered ERIs
// of a derivative plan (one block per quartet); raw_off[k * nq + q] and
// raw_off[(4 + k) * nq + q] locate quartet q's ERIs with centre k raised and
// lowered; qinfo[q] packs l (4 bits per centre) and the pure flags (bit 16 + k).
void launch_deriv1_combine(const double* raw, const long long* raw_off, const long long* out_off, const int* qinfo,
                           long long nq, long long total, double* out, const L2gDerivTab& tab, hipStream_t s);

// l2g_generic.hip: the table-driven kernel for classes without a generated
// register-resident kernel (f shells); modes ERI store, Fock digest, Schwarz
enum { kGenericEri = 0, kGenericFock = 1, kGenericSchwarz = 2 };
void generic_init(const double* sh);  // sh[4][7][10] solid-harmonic coefficients
void launch_generic(int mode, const L2gEriArgs& A, const L2gGenericClass& C, const int4* prog, const int* idx,
                    int grid, hipStream_t s);
inline int lmax_supported() { return std::max(l2g_gen_lmax, l2g_gen_generic_lmax); }

// l2g_enum.hip: the screened unique-quartet list of a Fock build on the GPU.
// Sort key: Fock kernel class (8 bits; the last value flags a class without
// a kernel) above 13 bits of decreasing primitive-loop length.
constexpr int kEnumClasses = 256;
constexpr int kEnumKeyBits = 21;
struct FockEnumArgs {
  const int* ps1;               // significant pairs in (s1, s2) order, s2 <= s1
  const int* ps2;
  const double* pK;             // Schwarz factor of each pair
  const double* pS;             // FP32-selection factor of each pair
  const int* pid;               // canonical shell-pair id (kernel pair data)
  const unsigned char* pswap;   // 1 when the canonical orientation is (s2, s1)
  const int* l;                 // per shell
  const int* pure;
  const int* bf1st;
  const int* np;                // primitive pairs per canonical pair id
  const double* Dn;             // n x n shell-block max norms of D
  long long n, npairs;
  const int* clsmap;            // [4][4][4][4] canonical class -> Fock kernel class (FP64)
  const int* cls32map;          // -> FP32 kernel class or -1
  double precision, needed, fp32_thr, hrr_factor;
};
struct FockEnumOut {
  int *bra, *ket, *meta;
  double *lnp, *scale;
  int4* bf;
  int2* pr;
  unsigned *key, *idx;
  unsigned long long *cls_count, *cls_nprim;
};
// the sorted items (class-major, decreasing depth) of the last enumeration
struct FockEnumState {
  // count pass: the triangular range split into nblocks blocks of tchunk
  // indices, with the surviving quartets of each
  long long total_t = 0, tchunk = 0;
  int nblocks = 0;
  std::vector<long long> block_count;
  long long n = 0;  // items of the filled window (after fock_split_device: pieces)
  std::vector<unsigned long long> cls_count, cls_nprim;  // kEnumClasses classes
  std::vector<unsigned long long> cls_pieces;            // per class, after the split
  DevBuf counts, base, perm, key, bra, ket, meta, lnp, scale, bf, pr;
  struct Items {
    int *bra, *ket, *meta;
    double *lnp, *scale;
    int4* bf;
    int2* pr;
  };
  Items items() {
    return Items{bra.as<int>(), ket.as<int>(), meta.as<int>(), lnp.as<double>(), scale.as<double>(), bf.as<int4>(),
                 pr.as<int2>()};
  }
};
void fock_enumerate_device(const FockEnumArgs& E, FockEnumState& S, hipStream_t st);
// the same in two steps: the count pass over the whole triangular range, then
// the items of count-pass blocks [blk0, blk1) (a window), sorted
void fock_count_device(const FockEnumArgs& E, FockEnumState& S, hipStream_t st);
void fock_fill_device(const FockEnumArgs& E, FockEnumState& S, int blk0, int blk1, hipStream_t st);
// Splits the enumerated items by primitive range, chunk[c] primitives at most
// per piece of class c (kEnumClasses entries); pieces stay class-major.
void fock_split_device(FockEnumState& S, const std::vector<int>& chunk, hipStream_t st);


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

  // Precomputed shell pairs (Engine::compute2 with ShellPair arguments,
  // engine.impl.h:1249-1285): the caller's ln_scr and nonsph_screen_fac of
  // every primitive pair of the ordered pair (i, j), p1-major; a primitive
  // pair whose ln_scr is -infinity (or NaN) is not in the caller's list.
  // When any pair is given, every pair must be.
  struct Given {
    const double* ln_scr;
    const double* nonsph;  // may be null (zeros)
  };
  std::unordered_map<uint64_t, Given> given;

  PairBuilder(const l2g_basis_s& b, double lp, int m) : B(b), ln_prec(lp), method(m) {}

  int get(int i, int j) {
    const uint64_t key = (uint64_t)(uint32_t)i << 32 | (uint32_t)j;
    auto it = id.find(key);
    if (it != id.end()) return it->second;
    const std::vector<double>* fac = nullptr;
    bool fac_t = false;
    if (given.empty() && (method == L2G_SCREEN_SCHWARZ || method == L2G_SCREEN_SCHWARZINF)) {
      auto f = pfac.find(key);
      if (f == pfac.end()) {
        f = pfac.find((uint64_t)(uint32_t)j << 32 | (uint32_t)i);
        fac_t = true;
      }
      if (f == pfac.end()) throw L2gError(L2G_E_STATE, "no primitive Schwarz factors for a shell pair");
      fac = f->second;
    }
    const Given* gv = nullptr;
    bool gv_t = false;
    if (!given.empty()) {
      auto g = given.find(key);
      if (g == given.end()) {
        g = given.find((uint64_t)(uint32_t)j << 32 | (uint32_t)i);
        gv_t = true;
      }
      if (g == given.end()) throw std::invalid_argument("no precomputed primitive-pair data for a shell pair");
      gv = &g->second;
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
        double nonsph = 0;
        if (gv) {
          const size_t k = gv_t ? (size_t)p2 * n1 + p1 : (size_t)p1 * n2 + p2;
          ln_scr = gv->ln_scr[k];
          if (!(ln_scr > -std::numeric_limits<double>::infinity())) continue;
          nonsph = gv->nonsph ? gv->nonsph[k] : 0.0;
        } else if (fac) {
          const double f = fac_t ? (*fac)[(size_t)p2 * n1 + p1] : (*fac)[(size_t)p1 * n2 + p2];
          ln_scr = std::log((double)(n1 * n2) * f) + B.lnc[B.off[i] + p1] + B.lnc[B.off[j] + p2];
          if (ln_scr < ln_prec) continue;
        }
        if (!gv && method == L2G_SCREEN_ORIGINAL && ln_scr < ln_prec) continue;
        double Pp[3];
        for (int k = 0; k < 3; ++k) Pp[k] = (AB2 == 0.0) ? A[k] : (a1 * A[k] + a2 * Bo[k]) * oog;
        if (!gv && method == L2G_SCREEN_CONSERVATIVE) {
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

// l2g_quartet.cpp: one shell quartet per call.  The staged input block of a
// call (byte offsets; np primitive pairs of both shell pairs together), then
// one double for the screened-out flag and the integrals.
inline constexpr size_t align32(size_t x) { return (x + 31) & ~size_t(31); }
struct QuartetLayout {
  size_t ints, lls, ab, P, PA, aux, ns, total;
  constexpr explicit QuartetLayout(size_t np)
      : ints(0),                                      // bra, ket, meta, pair_first[2], pair_np[2]
        lls(align32(8 * sizeof(int))),                // out_off, qidx
        ab(align32(lls + 2 * sizeof(long long))),     // AB of both pairs
        P(align32(ab + 6 * sizeof(double))),
        PA(align32(P + np * 4 * sizeof(double))),
        aux(align32(PA + np * 4 * sizeof(double))),
        ns(align32(aux + np * 2 * sizeof(double))),
        total(align32(ns + np * sizeof(double))) {}
};
constexpr int kQuartetBlock = 128;  // threads of the one block (the kernels' launch bound)

// Where one quartet evaluation works: the device block (inputs at most
// in_bytes, then the flag slot and out_doubles integrals), the lane
// scratch (ylane doubles per lane, kQuartetBlock lanes) and host staging
// of the inputs (in_bytes + one double) and of the flag and integrals.
struct QuartetBuffers {
  int max_l = 0, max_nprim = 0;
  unsigned char* d = nullptr;
  double* ybuf = nullptr;
  int64_t ylane = 0;
  unsigned char* hin = nullptr;
  double* hout = nullptr;
  size_t in_bytes = 0, out_doubles = 0;
  const char* owner = "quartet engine";  // names the limits in error messages
};
// lane scratch per lane of the largest register-kernel class with l <= max_l
int64_t quartet_ylane(int max_l);
void quartet_compute_impl(l2g_context_s* ctx, const QuartetBuffers& qb, const int* l, const int* pure,
                          const int* nprim, const double* alpha, const double* coeff, const double* origin,
                          const double* ln_scr, const double* nonsph, double precision, int screening, double scale,
                          int uniform_cart, int oper, double omega, double* out, int8_t* nonzero);
// l2g_noalloc.hip: resolves this device's copy of the no-allocation
// entry point's scratch array (called by l2g_init with the device current)
void noalloc_setup(int device);

}  // namespace l2g_impl
