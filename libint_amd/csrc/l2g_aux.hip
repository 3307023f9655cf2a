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
                                                              double* __restrict__ g, long long N) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * N) return;
  const long long i = idx / N, j = idx - i * N;
  if (j < i) return;
  const long long a = i * N + j, b = j * N + i;
  const double Jv = 0.25 * (jg[a] + jg[b]);
  const double Kv = 0.125 * (kg[a] + kg[b]);
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
// deriv_order = 1, include/libint2/engine.h:779-781).  A block per quartet;
// its threads stride over the 12 x nout outputs.
__global__ void __launch_bounds__(256) l2g_deriv1_combine_kernel(const double* __restrict__ raw,
                                                                 const long long* __restrict__ raw_off,
                                                                 const long long* __restrict__ out_off,
                                                                 const int* __restrict__ qinfo, long long nq,
                                                                 long long total, double* __restrict__ out,
                                                                 l2g_impl::L2gDerivTab tab) {
  const long long q = blockIdx.x;
  if (q >= nq) return;
  const int info = qinfo[q];
  int l[4], pure[4], nc[4], no[4];
  for (int k = 0; k < 4; ++k) {
    l[k] = (info >> (4 * k)) & 15;
    pure[k] = (info >> (16 + k)) & 1;
    nc[k] = (l[k] + 1) * (l[k] + 2) / 2;
    no[k] = pure[k] ? 2 * l[k] + 1 : nc[k];
  }
  const int nout = no[0] * no[1] * no[2] * no[3];
  const long long obase = out_off[q];
  for (int idx = threadIdx.x; idx < 12 * nout; idx += blockDim.x) {
    const int d = idx / nout;
    int f = idx - d * nout;
    const int k = d / 3, ax = d - 3 * k;
    int fo[4];
    for (int j = 3; j >= 0; --j) {
      fo[j] = f % no[j];
      f /= no[j];
    }
    const double* rb = raw + raw_off[(long long)k * nq + q];
    const double* lb = raw + raw_off[(long long)(4 + k) * nq + q];
    int dp[4], dm[4];
    for (int j = 0; j < 4; ++j) dp[j] = dm[j] = nc[j];
    dp[k] = (l[k] + 2) * (l[k] + 3) / 2;
    dm[k] = l[k] * (l[k] + 1) / 2;
    int n[4];
    for (int j = 0; j < 4; ++j) n[j] = pure[j] ? tab.nnz[l[j]][fo[j]] : 1;
    double sum = 0.0;
    for (int a0 = 0; a0 < n[0]; ++a0)
      for (int a1 = 0; a1 < n[1]; ++a1)
        for (int a2 = 0; a2 < n[2]; ++a2)
          for (int a3 = 0; a3 < n[3]; ++a3) {
            const int a[4] = {a0, a1, a2, a3};
            int c[4];
            double w = 1.0;
            for (int j = 0; j < 4; ++j) {
              if (pure[j]) {
                c[j] = tab.cidx[l[j]][fo[j]][a[j]];
                w *= tab.coef[l[j]][fo[j]][a[j]];
              } else {
                c[j] = fo[j];
              }
            }
            // exponents of component c[k] of shell k
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
            sum += w * v;
          }
    out[(long long)d * total + obase + idx - d * nout] = sum;
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

void launch_fock_finish(double* jg, double* kg, double* g, long long N, hipStream_t s) {
  if (N <= 0) return;
  const long long n = N * N;
  const int block = 256;
  const long long grid = (n + block - 1) / block;
  hipLaunchKernelGGL(l2g_fock_finish_kernel, dim3((unsigned)grid), dim3(block), 0, s, jg, kg, g, N);
  LAUNCHCHK("l2g_fock_finish_kernel");
}

}  // namespace l2g_impl
