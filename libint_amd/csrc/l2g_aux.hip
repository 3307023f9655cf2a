// Small hand-written kernels used by the host code: the Fock finishing step.
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

namespace l2g_impl {

void launch_fock_finish(double* jg, double* kg, double* g, long long N, hipStream_t s) {
  if (N <= 0) return;
  const long long n = N * N;
  const int block = 256;
  const long long grid = (n + block - 1) / block;
  hipLaunchKernelGGL(l2g_fock_finish_kernel, dim3((unsigned)grid), dim3(block), 0, s, jg, kg, g, N);
  LAUNCHCHK("l2g_fock_finish_kernel");
}

}  // namespace l2g_impl
