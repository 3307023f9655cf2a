// Screened Fock quartet list on the GPU: the permutationally-unique quartet
// loop of compute_2body_fock (tests/hartree-fock/hartree-fock++.cc:1628-1696)
// over the significant shell pairs, which are listed in (s1, s2)
// lexicographic order.  In that order the unique quartets are exactly the
// pairs of pairs (p1, p2) with p2 <= p1, so the loop is one triangular index
// t -> (p1, p2) that every thread decodes on its own.
//
// Pass 1 counts the surviving quartets of each block's contiguous t range;
// pass 2 writes them (screening Dnorm1234 K12 K34 >= precision, per-quartet
// precision precision / Dnorm1234, degeneracy, canonical shell order and
// kernel class as the host enumeration in l2g_fock.cpp) at the block's
// offset, with a sort key (class, decreasing primitive-loop length);
// a radix sort by that key then groups the classes and orders each by depth.
#include <hipcub/hipcub.hpp>

#include "l2g_internal.h"

using namespace l2g_impl;

namespace {

constexpr int kEnumThreads = 256;

__device__ __forceinline__ void decode_tri(long long t, long long& p1, long long& p2) {
  long long a = (long long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (a * (a + 1) / 2 > t) --a;
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  p1 = a;
  p2 = t - a * (a + 1) / 2;
}

struct Quartet {
  int s1, s2, s3, s4;
  double Dn1234;
};

// screening of one unique quartet; true when it survives
__device__ __forceinline__ bool screen(const l2g_impl::FockEnumArgs& E, long long p1, long long p2, Quartet& q) {
  q.s1 = E.ps1[p1], q.s2 = E.ps2[p1], q.s3 = E.ps1[p2], q.s4 = E.ps2[p2];
  const long long n = E.n;
  const double* Dn = E.Dn;
  const double d = fmax(fmax(fmax(Dn[q.s1 * n + q.s2], Dn[q.s3 * n + q.s4]), fmax(Dn[q.s1 * n + q.s3], Dn[q.s2 * n + q.s3])),
                        fmax(Dn[q.s1 * n + q.s4], Dn[q.s2 * n + q.s4]));
  q.Dn1234 = d;
  return !(d * E.pK[p1] * E.pK[p2] < E.precision);
}

__global__ void __launch_bounds__(kEnumThreads) enum_count(const l2g_impl::FockEnumArgs E, long long total,
                                                              long long chunk, long long* __restrict__ counts) {
  __shared__ long long red[kEnumThreads / 64];
  const long long t0 = (long long)blockIdx.x * chunk, t1 = min(total, t0 + chunk);
  long long c = 0;
  for (long long t = t0 + threadIdx.x; t < t1; t += kEnumThreads) {
    long long p1, p2;
    decode_tri(t, p1, p2);
    Quartet q;
    c += screen(E, p1, p2, q) ? 1 : 0;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long s = 0;
    for (int i = 0; i < kEnumThreads / 64; ++i) s += red[i];
    counts[blockIdx.x] = s;
  }
}

// fills the items of blocks [block0, block0 + gridDim.x) of the count pass's
// partition; base[k] is the first item of block block0 + k within the window
__global__ void __launch_bounds__(kEnumThreads) enum_fill(const l2g_impl::FockEnumArgs E, long long total,
                                                             long long chunk, long long block0,
                                                             const long long* __restrict__ base,
                                                             l2g_impl::FockEnumOut O) {
  __shared__ unsigned int pos;
  __shared__ unsigned int hist[l2g_impl::kEnumClasses];
  __shared__ unsigned long long whist[l2g_impl::kEnumClasses];
  if (threadIdx.x == 0) pos = 0;
  for (int c = threadIdx.x; c < l2g_impl::kEnumClasses; c += kEnumThreads) hist[c] = 0, whist[c] = 0;
  __syncthreads();
  const long long t0 = (block0 + blockIdx.x) * chunk, t1 = min(total, t0 + chunk);
  const long long b0 = base[blockIdx.x];
  for (long long t = t0 + threadIdx.x; t < t1; t += kEnumThreads) {
    long long p1, p2;
    decode_tri(t, p1, p2);
    Quartet q;
    if (!screen(E, p1, p2, q)) continue;
    const int s1 = q.s1, s2 = q.s2, s3 = q.s3, s4 = q.s4;
    const double qprec = q.Dn1234 != 0.0 ? E.precision / q.Dn1234 : E.needed;
    const int deg = (s1 == s2 ? 1 : 2) * (s3 == s4 ? 1 : 2) * (s1 == s3 ? (s2 == s4 ? 1 : 2) : 2);
    const int l1 = E.l[s1], l2 = E.l[s2], l3 = E.l[s3], l4 = E.l[s4];
    const bool stb = E.pswap[p1] != 0, stk = E.pswap[p2] != 0, sbk = (l1 + l2) > (l3 + l4);
    const int b1 = sbk ? (stk ? s4 : s3) : (stb ? s2 : s1);
    const int b2 = sbk ? (stk ? s3 : s4) : (stb ? s1 : s2);
    const int k1 = sbk ? (stb ? s2 : s1) : (stk ? s4 : s3);
    const int k2 = sbk ? (stb ? s1 : s2) : (stk ? s3 : s4);
    const int bp = sbk ? E.pid[p2] : E.pid[p1], kp = sbk ? E.pid[p1] : E.pid[p2];
    const int lk = ((E.l[b1] * 4 + E.l[b2]) * 4 + E.l[k1]) * 4 + E.l[k2];
    // every class up to the compiled l has a kernel; the last class flags none
    int cls = E.clsmap[lk];
    if (E.fp32_thr > 0 && E.pK[p1] * E.pK[p2] < E.fp32_thr && E.pS[p1] * E.pS[p2] < E.hrr_factor * E.fp32_thr &&
        E.cls32map[lk] >= 0)
      cls = E.cls32map[lk];
    if (cls < 0 || cls >= l2g_impl::kEnumClasses - 1) cls = l2g_impl::kEnumClasses - 1;
    auto pb = [&](int sh) { return (E.pure[sh] && E.l[sh] > 0) ? 1 : 0; };
    const int w = E.np[bp] * E.np[kp];
    const long long i = b0 + atomicAdd(&pos, 1u);
    O.bra[i] = bp;
    O.ket[i] = kp;
    O.meta[i] = (pb(b1) << 3) | (pb(b2) << 4) | (pb(k1) << 5) | (pb(k2) << 6);
    O.lnp[i] = qprec > 0 ? log(qprec) : -1.7976931348623157e308;
    O.scale[i] = (double)deg;
    O.bf[i] = int4{E.bf1st[b1], E.bf1st[b2], E.bf1st[k1], E.bf1st[k2]};
    O.pr[i] = int2{0, w};
    O.key[i] = ((unsigned)cls << 13) | (unsigned)(8191 - min(w, 8191));  // kEnumKeyBits bits
    O.idx[i] = (unsigned)i;
    atomicAdd(&hist[cls], 1u);
    atomicAdd(&whist[cls], (unsigned long long)w);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < l2g_impl::kEnumClasses; c += kEnumThreads)
    if (hist[c]) {
      atomicAdd(&O.cls_count[c], (unsigned long long)hist[c]);
      atomicAdd(&O.cls_nprim[c], whist[c]);
    }
}

template <class T>
__global__ void gather_kernel(const T* __restrict__ in, const unsigned* __restrict__ perm, T* __restrict__ out,
                              long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[perm[i]];
}

template <class T>
void gather(const T* in, const unsigned* perm, T* out, long long n, hipStream_t s) {
  if (n == 0) return;
  const int block = 256;
  hipLaunchKernelGGL(gather_kernel<T>, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, s, in, perm, out, n);
  LAUNCHCHK("gather_kernel");
}

// Primitive-range split of the sorted items: an item of class c whose
// primitive loop is longer than chunk[c] becomes ceil(len / chunk[c])
// consecutive pieces of near-equal length.
__global__ void __launch_bounds__(kEnumThreads) split_count(long long n, const unsigned* __restrict__ key,
                                                               const int2* __restrict__ pr,
                                                               const int* __restrict__ chunk,
                                                               unsigned* __restrict__ pieces,
                                                               unsigned long long* __restrict__ cls_pieces) {
  __shared__ unsigned int hist[l2g_impl::kEnumClasses];
  for (int c = threadIdx.x; c < l2g_impl::kEnumClasses; c += kEnumThreads) hist[c] = 0;
  __syncthreads();
  const long long i = (long long)blockIdx.x * kEnumThreads + threadIdx.x;
  if (i < n) {
    const int c = (int)(key[i] >> 13);
    const int len = pr[i].y - pr[i].x, ch = chunk[c];
    const unsigned np = len > ch ? (unsigned)((len + ch - 1) / ch) : 1u;
    pieces[i] = np;
    atomicAdd(&hist[c], np);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < l2g_impl::kEnumClasses; c += kEnumThreads)
    if (hist[c]) atomicAdd(&cls_pieces[c], (unsigned long long)hist[c]);
}

__global__ void __launch_bounds__(kEnumThreads) split_fill(long long n, const unsigned* __restrict__ pieces,
                                                              const unsigned long long* __restrict__ at,
                                                              l2g_impl::FockEnumState::Items in,
                                                              l2g_impl::FockEnumState::Items out) {
  const long long i = (long long)blockIdx.x * kEnumThreads + threadIdx.x;
  if (i >= n) return;
  const unsigned np = pieces[i];
  const unsigned long long o = at[i];
  const int b = in.pr[i].x, len = in.pr[i].y - b;
  for (unsigned k = 0; k < np; ++k) {
    const unsigned long long j = o + k;
    out.bra[j] = in.bra[i];
    out.ket[j] = in.ket[i];
    out.meta[j] = in.meta[i];
    out.lnp[j] = in.lnp[i];
    out.scale[j] = in.scale[i];
    out.bf[j] = in.bf[i];
    out.pr[j] = int2{b + (int)((long long)len * k / np), b + (int)((long long)len * (k + 1) / np)};
  }
}

}  // namespace

namespace l2g_impl {

void fock_split_device(FockEnumState& S, const std::vector<int>& chunk, hipStream_t st) {
  const long long n = S.n;
  S.cls_pieces.assign(kEnumClasses, 0ull);
  if (n == 0) return;
  if ((int)chunk.size() != kEnumClasses) throw L2gError(L2G_E_STATE, "split: one chunk per class expected");
  LAUNCHCHK("work before the item split");
  DevBuf d_chunk, pieces, at, cls, tmp;
  d_chunk.alloc(kEnumClasses * sizeof(int));
  HIPCHK(hipMemcpyAsync(d_chunk.p, chunk.data(), kEnumClasses * sizeof(int), hipMemcpyHostToDevice, st));
  pieces.alloc((size_t)n * sizeof(unsigned));
  at.alloc((size_t)n * sizeof(unsigned long long));
  cls.alloc(kEnumClasses * sizeof(unsigned long long));
  HIPCHK(hipMemsetAsync(cls.p, 0, kEnumClasses * sizeof(unsigned long long), st));
  const unsigned nb = (unsigned)((n + kEnumThreads - 1) / kEnumThreads);
  hipLaunchKernelGGL(split_count, dim3(nb), dim3(kEnumThreads), 0, st, n, S.key.as<unsigned>(), S.pr.as<int2>(),
                     d_chunk.as<int>(), pieces.as<unsigned>(), cls.as<unsigned long long>());
  LAUNCHCHK("split_count");
  size_t tmp_bytes = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, pieces.as<unsigned>(), at.as<unsigned long long>(),
                                          (int)n, st));
  tmp.alloc(std::max<size_t>(tmp_bytes, 16));
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, pieces.as<unsigned>(), at.as<unsigned long long>(),
                                          (int)n, st));
  HIPCHK(hipMemcpyAsync(S.cls_pieces.data(), cls.p, kEnumClasses * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  unsigned long long total = 0;
  for (int c = 0; c < kEnumClasses; ++c) {
    if (S.cls_pieces[c] < S.cls_count[c])
      throw L2gError(L2G_E_STATE, "split: class " + std::to_string(c) + " has " + std::to_string(S.cls_pieces[c]) +
                                      " pieces for " + std::to_string(S.cls_count[c]) + " items");
    total += S.cls_pieces[c];
  }
  if (total >= (1ull << 31)) throw L2gError(L2G_E_INVALID, "more than 2^31 work items in one Fock build");
  if (total == (unsigned long long)n) return;  // nothing to split
  FockEnumState::Items in = S.items();
  DevBuf bra, ket, meta, lnp, scale, bf, pr;
  bra.alloc(total * sizeof(int));
  ket.alloc(total * sizeof(int));
  meta.alloc(total * sizeof(int));
  lnp.alloc(total * sizeof(double));
  scale.alloc(total * sizeof(double));
  bf.alloc(total * sizeof(int4));
  pr.alloc(total * sizeof(int2));
  FockEnumState::Items out{bra.as<int>(), ket.as<int>(), meta.as<int>(), lnp.as<double>(), scale.as<double>(),
                           bf.as<int4>(), pr.as<int2>()};
  hipLaunchKernelGGL(split_fill, dim3(nb), dim3(kEnumThreads), 0, st, n, pieces.as<unsigned>(),
                     at.as<unsigned long long>(), in, out);
  LAUNCHCHK("split_fill");
  HIPCHK(hipStreamSynchronize(st));  // the inputs are freed on return
  std::swap(S.bra, bra), std::swap(S.ket, ket), std::swap(S.meta, meta), std::swap(S.lnp, lnp);
  std::swap(S.scale, scale), std::swap(S.bf, bf), std::swap(S.pr, pr);
  S.n = (long long)total;
}

void fock_count_device(const FockEnumArgs& E, FockEnumState& S, hipStream_t st) {
  const long long np = E.npairs;
  S.total_t = np * (np + 1) / 2;
  S.block_count.clear();
  S.nblocks = 0;
  S.tchunk = 0;
  if (S.total_t == 0) return;
  S.nblocks = (int)std::min<long long>(8192, (S.total_t + kEnumThreads * 64 - 1) / (kEnumThreads * 64));
  S.tchunk = (S.total_t + S.nblocks - 1) / S.nblocks;
  S.counts.alloc((size_t)S.nblocks * sizeof(long long));
  hipLaunchKernelGGL(enum_count, dim3(S.nblocks), dim3(kEnumThreads), 0, st, E, S.total_t, S.tchunk,
                     S.counts.as<long long>());
  LAUNCHCHK("enum_count");
  S.block_count.resize(S.nblocks);
  HIPCHK(hipMemcpyAsync(S.block_count.data(), S.counts.p, (size_t)S.nblocks * sizeof(long long),
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
}

void fock_fill_device(const FockEnumArgs& E, FockEnumState& S, int blk0, int blk1, hipStream_t st) {
  S.n = 0;
  S.cls_count.assign(kEnumClasses, 0ull);
  S.cls_nprim.assign(kEnumClasses, 0ull);
  if (blk1 <= blk0) return;
  const int nb = blk1 - blk0;
  std::vector<long long> base(nb);
  long long n = 0;
  for (int b = 0; b < nb; ++b) {
    base[b] = n;
    n += S.block_count[blk0 + b];
  }
  S.n = n;
  if (n == 0) return;
  if (n >= (1ll << 31)) throw L2gError(L2G_E_INVALID, "more than 2^31 screened quartets in one window of a Fock build");
  S.base.alloc((size_t)nb * sizeof(long long));
  HIPCHK(hipMemcpyAsync(S.base.p, base.data(), (size_t)nb * sizeof(long long), hipMemcpyHostToDevice, st));
  // unsorted items, then sorted
  const size_t sz = (size_t)n;
  DevBuf u_bra, u_ket, u_meta, u_lnp, u_scale, u_bf, u_pr, u_key, u_idx, stats;
  DevBuf& s_key = S.key;
  u_bra.alloc(sz * sizeof(int));
  u_ket.alloc(sz * sizeof(int));
  u_meta.alloc(sz * sizeof(int));
  u_lnp.alloc(sz * sizeof(double));
  u_scale.alloc(sz * sizeof(double));
  u_bf.alloc(sz * sizeof(int4));
  u_pr.alloc(sz * sizeof(int2));
  u_key.alloc(sz * sizeof(unsigned));
  u_idx.alloc(sz * sizeof(unsigned));
  s_key.alloc(sz * sizeof(unsigned));
  stats.alloc(2 * kEnumClasses * sizeof(unsigned long long));
  HIPCHK(hipMemsetAsync(stats.p, 0, 2 * kEnumClasses * sizeof(unsigned long long), st));
  FockEnumOut O;
  O.bra = u_bra.as<int>();
  O.ket = u_ket.as<int>();
  O.meta = u_meta.as<int>();
  O.lnp = u_lnp.as<double>();
  O.scale = u_scale.as<double>();
  O.bf = u_bf.as<int4>();
  O.pr = u_pr.as<int2>();
  O.key = u_key.as<unsigned>();
  O.idx = u_idx.as<unsigned>();
  O.cls_count = stats.as<unsigned long long>();
  O.cls_nprim = stats.as<unsigned long long>() + kEnumClasses;
  hipLaunchKernelGGL(enum_fill, dim3(nb), dim3(kEnumThreads), 0, st, E, S.total_t, S.tchunk, (long long)blk0,
                     S.base.as<long long>(), O);
  LAUNCHCHK("enum_fill");
  // sort the item indices by (class, decreasing depth)
  S.perm.alloc(sz * sizeof(unsigned));
  size_t tmp_bytes = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, O.key, s_key.as<unsigned>(), O.idx,
                                            S.perm.as<unsigned>(), (int)n, 0, kEnumKeyBits, st));
  DevBuf tmp;
  tmp.alloc(std::max<size_t>(tmp_bytes, 16));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, O.key, s_key.as<unsigned>(), O.idx,
                                            S.perm.as<unsigned>(), (int)n, 0, kEnumKeyBits, st));
  const unsigned* perm = S.perm.as<unsigned>();
  S.bra.alloc(sz * sizeof(int));
  S.ket.alloc(sz * sizeof(int));
  S.meta.alloc(sz * sizeof(int));
  S.lnp.alloc(sz * sizeof(double));
  S.scale.alloc(sz * sizeof(double));
  S.bf.alloc(sz * sizeof(int4));
  S.pr.alloc(sz * sizeof(int2));
  gather(O.bra, perm, S.bra.as<int>(), n, st);
  gather(O.ket, perm, S.ket.as<int>(), n, st);
  gather(O.meta, perm, S.meta.as<int>(), n, st);
  gather(O.lnp, perm, S.lnp.as<double>(), n, st);
  gather(O.scale, perm, S.scale.as<double>(), n, st);
  gather(O.bf, perm, S.bf.as<int4>(), n, st);
  gather(O.pr, perm, S.pr.as<int2>(), n, st);
  std::vector<unsigned long long> h(2 * kEnumClasses);
  HIPCHK(hipMemcpyAsync(h.data(), stats.p, 2 * kEnumClasses * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));  // the temporaries are freed on return
  for (int c = 0; c < kEnumClasses; ++c) S.cls_count[c] = h[c], S.cls_nprim[c] = h[kEnumClasses + c];
  if (S.cls_count[kEnumClasses - 1]) throw L2gError(L2G_E_LMAX, "no Fock kernel for an angular momentum class");
}

void fock_enumerate_device(const FockEnumArgs& E, FockEnumState& S, hipStream_t st) {
  fock_count_device(E, S, st);
  fock_fill_device(E, S, 0, S.nblocks, st);
}

}  // namespace l2g_impl
