// Small hand-written kernels used by the host code: the Fock finishing step
// and the combination step of first-derivative integrals.
#include "l2g_internal.h"

// In place, for one request block of the Fock arenas: J = (jg + jg^T)/4 into
// jg, K = (kg + kg^T)/8 into kg, G = 2J - K into the density block.  The
// digest kernels add each unique quartet's integrals once with its
// degeneracy factor and without mirroring, which leaves this
// symmetrisation, as in compute_2body_fock
// (tests/hartree-fock/hartree-fock++.cc:1709-1771).
// Thread (i, j) with i <= j owns both mirrored elements, so no element is
// read after another thread wrote it.
__global__ void __launch_bounds__(256) l2g_fock_finish_kernel(double* __restrict__ jg, double* __restrict__ kg,
                                                              double* __restrict__ g, long long N, long long hoff) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * N) return;
  const long long i = idx / N, j = idx - i * N;
  if (j < i) return;
  const long long a = i * N + j, b = j * N + i;
  double ja, jb, ka, kb;
  if (hoff) {  // fixed-point accumulators: low words in place, high words hoff further on
    const unsigned long long* jw = reinterpret_cast<const unsigned long long*>(jg);
    const unsigned long long* kw = reinterpret_cast<const unsigned long long*>(kg);
    ja = l2g_fx_value(jw[a + hoff], jw[a]);
    jb = l2g_fx_value(jw[b + hoff], jw[b]);
    ka = l2g_fx_value(kw[a + hoff], kw[a]);
    kb = l2g_fx_value(kw[b + hoff], kw[b]);
  } else {
    ja = jg[a], jb = jg[b], ka = kg[a], kb = kg[b];
  }
  const double Jv = 0.25 * (ja + jb);
  const double Kv = 0.125 * (ka + kb);
  jg[a] = Jv;
  jg[b] = Jv;
  kg[a] = Kv;
  kg[b] = Kv;
  g[a] = 2.0 * Jv - Kv;
  g[b] = 2.0 * Jv - Kv;
}

// Cartesian component index of x^a y^b z^c in libint's order (x descending,
// then y descending): (l - a)(l - a + 1)/2 + (l - a - b).
__device__ __forceinline__ int l2g_cart_index(int l, int a, int b) { return (l - a) * (l - a + 1) / 2 + (l - a - b); }

// First derivatives of (ab|cd) with respect to the centre k of one shell,
// d/dK_i x_i^n e^{-alpha r^2} = 2 alpha x_i^(n+1) e^{..} - n x_i^(n-1) e^{..}:
// the ERI with shell k raised (l+1, coefficients 2 alpha c) minus n times
// the ERI with shell k lowered (l-1), both Cartesian, then the solid-harmonic
// rows of the pure centres.  Derivative d = 3 k + i of quartet q goes to
// out[d * total + out_off[q] + f], f row-major over the four shells' output
// functions in caller order (the layout of Engine::results()[d] for
// deriv_order = 1, include/libint2/engine.h:779-781).  A block per quartet
// takes the 12 derivatives one at a time: the Cartesian block (two loads
// per element) into LDS, then one centre's solid-harmonic rows after the
// other between two LDS buffers, then the store.  Shells up to l = 2 (the
// raised ones up to 3): at most 6^4 Cartesian elements.
constexpr int kDerivMaxCart = 1296;

__global__ void __launch_bounds__(256) l2g_deriv1_combine_kernel(const double* __restrict__ raw,
                                                                 const long long* __restrict__ raw_off,
                                                                 const long long* __restrict__ out_off,
                                                                 const int* __restrict__ qinfo, long long nq,
                                                                 long long total, double* __restrict__ out,
                                                                 l2g_impl::L2gDerivTab tab) {
  __shared__ double buf[2][kDerivMaxCart];
  const long long q = blockIdx.x;
  if (q >= nq) return;
  const int info = qinfo[q];
  int l[4], pure[4], nc[4];
  for (int k = 0; k < 4; ++k) {
    l[k] = (info >> (4 * k)) & 15;
    pure[k] = (info >> (16 + k)) & 1;
    nc[k] = (l[k] + 1) * (l[k] + 2) / 2;
  }
  const int ncart = nc[0] * nc[1] * nc[2] * nc[3];
  const long long obase = out_off[q];
  for (int d = 0; d < 12; ++d) {
    const int k = d / 3, ax = d - 3 * k;
    const double* rb = raw + raw_off[(long long)k * nq + q];
    const double* lb = raw + raw_off[(long long)(4 + k) * nq + q];
    int dp[4], dm[4];
    for (int j = 0; j < 4; ++j) dp[j] = dm[j] = nc[j];
    dp[k] = (l[k] + 2) * (l[k] + 3) / 2;
    dm[k] = l[k] * (l[k] + 1) / 2;
    // Cartesian derivative block
    for (int idx = threadIdx.x; idx < ncart; idx += blockDim.x) {
      int c[4], f = idx;
      for (int j = 3; j >= 0; --j) {
        c[j] = f % nc[j];
        f /= nc[j];
      }
      int r = 0;
      while ((r + 1) * (r + 2) / 2 <= c[k]) ++r;
      int e[3];
      e[2] = c[k] - r * (r + 1) / 2;
      e[1] = r - e[2];
      e[0] = l[k] - r;
      int cu[4] = {c[0], c[1], c[2], c[3]};
      ++e[ax];
      cu[k] = l2g_cart_index(l[k] + 1, e[0], e[1]);
      --e[ax];
      double v = rb[((long long)(cu[0] * dp[1] + cu[1]) * dp[2] + cu[2]) * dp[3] + cu[3]];
      if (e[ax] > 0) {
        const int nx = e[ax];
        --e[ax];
        cu[k] = l2g_cart_index(l[k] - 1, e[0], e[1]);
        v -= nx * lb[((long long)(cu[0] * dm[1] + cu[1]) * dm[2] + cu[2]) * dm[3] + cu[3]];
      }
      buf[0][idx] = v;
    }
    __syncthreads();
    // solid-harmonic rows, one pure centre at a time
    int dims[4] = {nc[0], nc[1], nc[2], nc[3]};
    int cur = 0;
    for (int j = 0; j < 4; ++j) {
      if (!pure[j]) continue;
      const int nout_j = 2 * l[j] + 1;
      int inner = 1;
      for (int t = j + 1; t < 4; ++t) inner *= dims[t];
      int outer = 1;
      for (int t = 0; t < j; ++t) outer *= dims[t];
      const int n_new = outer * nout_j * inner;
      for (int idx = threadIdx.x; idx < n_new; idx += blockDim.x) {
        const int i_in = idx % inner, rest = idx / inner;
        const int m = rest % nout_j, o = rest / nout_j;
        double s = 0.0;
        for (int a = 0; a < tab.nnz[l[j]][m]; ++a)
          s += tab.coef[l[j]][m][a] * buf[cur][(o * dims[j] + tab.cidx[l[j]][m][a]) * inner + i_in];
        buf[cur ^ 1][idx] = s;
      }
      dims[j] = nout_j;
      cur ^= 1;
      __syncthreads();
    }
    const int nout = dims[0] * dims[1] * dims[2] * dims[3];
    double* dst = out + (long long)d * total + obase;
    for (int idx = threadIdx.x; idx < nout; idx += blockDim.x) dst[idx] = buf[cur][idx];
    __syncthreads();
  }
}

namespace l2g_impl {

void launch_deriv1_combine(const double* raw, const long long* raw_off, const long long* out_off, const int* qinfo,
                           long long nq, long long total, double* out, const L2gDerivTab& tab, hipStream_t s) {
  if (nq <= 0) return;
  if (nq > 0x7fffffffLL) throw std::invalid_argument("too many quartets for one derivative launch");
  hipLaunchKernelGGL(l2g_deriv1_combine_kernel, dim3((unsigned)nq), dim3(256), 0, s, raw, raw_off, out_off, qinfo, nq,
                     total, out, tab);
  LAUNCHCHK("l2g_deriv1_combine_kernel");
}

void launch_fock_finish(double* jg, double* kg, double* g, long long N, long long hoff, hipStream_t s) {
  if (N <= 0) return;
  const long long n = N * N;
  const int block = 256;
  const long long grid = (n + block - 1) / block;
  hipLaunchKernelGGL(l2g_fock_finish_kernel, dim3((unsigned)grid), dim3(block), 0, s, jg, kg, g, N, hoff);
  LAUNCHCHK("l2g_fock_finish_kernel");
}

}  // namespace l2g_impl
