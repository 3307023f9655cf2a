// The single-quartet evaluator for callers that cannot allocate: no
// hipMalloc, hipFree or pinned host allocation on any call.  Its device
// memory (the staged inputs, the screened-out flag, the integrals and the
// lane scratch of the class kernels) is the file-scope array `scratch`
// below, of fixed size, which each GPU the library runs on holds until the
// process ends.  l2g_init resolves the calling device's address of the
// array once (hipGetSymbolAddress with that device current); host code only
// ever hands kernels and copies that address.
//
// Calls on one device share its array.  They are serialised by a lock of
// the device, so calls from several threads of one process are safe and
// each leaves the array holding its own inputs and integrals; a kernel of
// the caller's that touched the array concurrently would race.
#include <mutex>

#include "l2g_internal.h"

using namespace l2g_impl;

namespace {

constexpr int kMaxL = 3;          // f shells
constexpr size_t kMaxYLane = 9216;  // lane scratch doubles per lane
constexpr size_t kInBytes = QuartetLayout(2 * (size_t)L2G_NOALLOC_MAX_NPRIM * L2G_NOALLOC_MAX_NPRIM).total;
constexpr size_t kOutDoubles = 10 * 10 * 10 * 10;  // (ff|ff) in Cartesian functions
constexpr size_t kYDoubles = kMaxYLane * kQuartetBlock;
// lane scratch first (256-byte aligned with the array), then the input
// block, the flag slot and the integrals
constexpr size_t kScratchDoubles = kYDoubles + kInBytes / sizeof(double) + 1 + kOutDoubles;
static_assert(kInBytes % sizeof(double) == 0, "input block is a whole number of doubles");

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
struct DeviceSlot {
  std::mutex lock;       // serialises calls on the device
  double* ptr = nullptr;  // the device's address of `scratch`
  std::string error;     // why resolving it failed
};
DeviceSlot g_slots[kMaxDevices];
std::mutex g_setup;

// host staging, one set per calling thread (no pinned memory)
alignas(64) thread_local unsigned char t_in[kInBytes + sizeof(double)];
alignas(64) thread_local double t_out[kOutDoubles + 1];

}  // namespace

namespace l2g_impl {

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_setup);
  DeviceSlot& s = g_slots[device];
  if (s.ptr) return;
  void* p = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (rc != hipSuccess || !p) {
    s.error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc);
    (void)hipGetLastError();
    return;
  }
  s.ptr = static_cast<double*>(p);
  s.error.clear();
}

}  // namespace l2g_impl

extern "C" {

int l2g_quartet_compute_noalloc(l2g_context ctx, const int* l, const int* pure, const int* nprim,
                                const double* alpha, const double* coeff, const double* origin, const double* ln_scr,
                                const double* nonsph, double precision, int screening, double scale, int uniform_cart,
                                int oper, double omega, double* out, int8_t* nonzero) {
  return guarded([&] {
    if (!ctx) throw L2gError(L2G_E_STATE, "null context");
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
      throw L2gError(L2G_E_STATE, "device index beyond the no-allocation entry point's table");
    DeviceSlot& s = g_slots[ctx->device];
    std::lock_guard<std::mutex> g(s.lock);
    if (!s.ptr)
      throw L2gError(L2G_E_STATE, "no-allocation scratch of device " + std::to_string(ctx->device) +
                                      " is not set up" + (s.error.empty() ? std::string() : ": " + s.error));
    const int64_t ylane = quartet_ylane(kMaxL);
    if (ylane > (int64_t)kMaxYLane)
      throw std::logic_error("no-allocation scratch holds " + std::to_string(kMaxYLane) +
                             " lane-scratch doubles per lane; the class kernels need " + std::to_string(ylane));
    QuartetBuffers qb;
    qb.max_l = std::min(kMaxL, lmax_supported());
    qb.max_nprim = L2G_NOALLOC_MAX_NPRIM;
    qb.ybuf = s.ptr;
    qb.ylane = kMaxYLane;
    qb.d = reinterpret_cast<unsigned char*>(s.ptr + kYDoubles);
    qb.in_bytes = kInBytes;
    qb.out_doubles = kOutDoubles;
    qb.hin = t_in;
    qb.hout = t_out;
    qb.owner = "no-allocation entry point";
    quartet_compute_impl(ctx, qb, l, pure, nprim, alpha, coeff, origin, ln_scr, nonsph, precision, screening, scale,
                         uniform_cart, oper, omega, out, nonzero);
  });
}

}  // extern "C"
