// Fock build (compute_2body_fock restated for the GPU) and the Fock-build
// request server of the C ABI declared in include/l2g.h.
#include <chrono>
#include <climits>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <numeric>

#include "l2g_internal.h"

using namespace l2g_impl;

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

// ---------------------------------------------------------------------------
// Fock setup: significant pairs, Schwarz factors and the shell-pair data of
// every significant pair, built once per basis and kept on the device
// ---------------------------------------------------------------------------
struct FockSetup {
  l2g_basis_s* B = nullptr;
  int method = L2G_SCREEN_SCHWARZINF;
  double max_engine_precision = 0;
  int n = 0, nbf = 0;
  std::vector<int> bf1st;
  std::vector<std::vector<int>> splist;  // splist[s1]: ascending s2 <= s1 of significant pairs
  std::vector<double> K, S;              // n x n Schwarz and FP32-selection factors, significant pairs only
  std::vector<int> pair_of;              // n x n pair id (canonical orientation), -1 if not significant
  // shell-pair data (PairBuilder layout) of the significant pairs
  std::vector<int> first, np;
  std::vector<double> AB;
  std::vector<double4> P, PA;
  std::vector<double2> aux;
  std::vector<double> ns;
};

// compute_shellpairs + compute_schwarz_ints (hartree-fock++.cc:1231-1436) and
// the primitive-pair factors of Schwarz/SchwarzInf screening (ShellPair::init,
// shell.h:1258-1328), with Schwarz kernels on the basis' context.
void fock_setup(FockSetup& f, l2g_basis_s* B, double pair_threshold, double max_engine_precision, int screening) {
  if (screening != L2G_SCREEN_ORIGINAL && screening != L2G_SCREEN_CONSERVATIVE && screening != L2G_SCREEN_SCHWARZ &&
      screening != L2G_SCREEN_SCHWARZINF)
    throw std::invalid_argument("unknown screening method");
  if (!(max_engine_precision > 0)) throw std::invalid_argument("max_engine_precision must be > 0");
  f.B = B;
  f.method = screening;
  f.max_engine_precision = max_engine_precision;
  const int n = B->nshell;
  f.n = n;
  f.bf1st.resize(n + 1);
  f.bf1st[0] = 0;
  for (int s = 0; s < n; ++s) f.bf1st[s + 1] = f.bf1st[s] + nfun(B->l[s], B->pure[s]);
  f.nbf = f.bf1st[n];
  // significant shell pairs: same centre, or overlap-block norm >= threshold
  f.splist.assign(n, {});
  parallel_for(n, [&](int s1) {
    for (int s2 = 0; s2 <= s1; ++s2) {
      const bool same = B->origin[3 * s1] == B->origin[3 * s2] && B->origin[3 * s1 + 1] == B->origin[3 * s2 + 1] &&
                        B->origin[3 * s1 + 2] == B->origin[3 * s2 + 2];
      if (same || overlap_norm(*B, s1, s2) >= pair_threshold) f.splist[s1].push_back(s2);
    }
  });
  std::vector<int32_t> pairs;
  for (int s1 = 0; s1 < n; ++s1)
    for (int s2 : f.splist[s1]) {
      pairs.push_back(s1);
      pairs.push_back(s2);
    }
  const int64_t np = (int64_t)pairs.size() / 2;
  // Schwarz factors of the significant pairs, no primitive screening
  std::vector<double> Kp(np), Sp(np);
  if (np > 0) schwarz_impl(B, np, pairs.data(), 0.0, 0, Kp.data(), Sp.data());
  f.K.assign((size_t)n * n, 0.0);
  f.S.assign((size_t)n * n, 0.0);
  for (int64_t i = 0; i < np; ++i) {
    const int a = pairs[2 * i], b = pairs[2 * i + 1];
    f.K[(size_t)a * n + b] = f.K[(size_t)b * n + a] = Kp[i];
    f.S[(size_t)a * n + b] = f.S[(size_t)b * n + a] = Sp[i];
  }
  // primitive-pair Schwarz factors from normalization-free primitives
  // (Shell::extract_primitive(p, false), shell.h:938-947), on the GPU
  std::vector<std::vector<double>> pfac;
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
    int rc = l2g_basis_create(B->ctx, (int)pl.size(), pl.data(), pp.data(), pn.data(), pa.data(), pc.data(), po.data(), &PB);
    if (rc) throw L2gError(rc, g_err);
    std::unique_ptr<l2g_basis_s, void (*)(l2g_basis_s*)> hold(PB, &basis_release);
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
    if (!kf.empty()) {
      rc = l2g_schwarz(PB, (int64_t)kf.size(), ppairs.data(), 0.0, screening == L2G_SCREEN_SCHWARZINF ? 0 : 1,
                       kf.data());
      if (rc) throw L2gError(rc, g_err);
    }
    pfac.resize(np);
    size_t k = 0;
    for (int64_t i = 0; i < np; ++i) {
      const int a = pairs[2 * i], b = pairs[2 * i + 1];
      pfac[i].assign(kf.begin() + k, kf.begin() + k + (size_t)B->nprim[a] * B->nprim[b]);
      k += (size_t)B->nprim[a] * B->nprim[b];
    }
  }
  // shell-pair data of every significant pair in canonical orientation
  PairBuilder pb(*B, std::log(max_engine_precision), screening);
  for (int64_t i = 0; i < (int64_t)pfac.size(); ++i)
    pb.pfac[(uint64_t)(uint32_t)pairs[2 * i] << 32 | (uint32_t)pairs[2 * i + 1]] = &pfac[i];
  f.pair_of.assign((size_t)n * n, -1);
  for (int64_t i = 0; i < np; ++i) {
    const int a = pairs[2 * i], b = pairs[2 * i + 1];
    const int id = B->swap_pair(a, b) ? pb.get(b, a) : pb.get(a, b);
    f.pair_of[(size_t)a * n + b] = f.pair_of[(size_t)b * n + a] = id;
  }
  f.first = std::move(pb.first);
  f.np = std::move(pb.np);
  f.AB = std::move(pb.AB);
  f.P = std::move(pb.P);
  f.PA = std::move(pb.PA);
  f.aux = std::move(pb.aux);
  f.ns = std::move(pb.ns);
}

// Work items of one Fock kernel class (struct of arrays, launch order).
struct FockItems {
  std::vector<int> bra, ket, meta, req;
  std::vector<double> lnp, scale;
  std::vector<int4> bf;
  std::vector<int2> pr;  // primitive-quartet range [x, y)
  size_t size() const { return bra.size(); }
  void append(const FockItems& o, size_t b = 0, size_t e = SIZE_MAX) {
    e = std::min(e, o.size());
    if (b >= e) return;
    bra.insert(bra.end(), o.bra.begin() + b, o.bra.begin() + e);
    ket.insert(ket.end(), o.ket.begin() + b, o.ket.begin() + e);
    meta.insert(meta.end(), o.meta.begin() + b, o.meta.begin() + e);
    lnp.insert(lnp.end(), o.lnp.begin() + b, o.lnp.begin() + e);
    scale.insert(scale.end(), o.scale.begin() + b, o.scale.begin() + e);
    bf.insert(bf.end(), o.bf.begin() + b, o.bf.begin() + e);
    pr.insert(pr.end(), o.pr.begin() + b, o.pr.begin() + e);
    if (!o.req.empty()) req.insert(req.end(), o.req.begin() + b, o.req.begin() + e);
  }
  void clear() {
    bra.clear(), ket.clear(), meta.clear(), req.clear(), lnp.clear(), scale.clear(), bf.clear(), pr.clear();
  }
};

// Fock kernel classes: generated FP64 classes, generated FP32 classes, then
// the generic kernel's classes (FP64).
int nfock_classes() { return l2g_gen_nfock + l2g_gen_nfock32 + l2g_gen_ngeneric; }
bool fock_is_generic(int c) { return c >= l2g_gen_nfock + l2g_gen_nfock32; }
const L2gGenericClass& fock_gclass(int c) { return l2g_gen_generic[c - l2g_gen_nfock - l2g_gen_nfock32]; }
const L2gClassInfo& fock_info(int c) {
  return c < l2g_gen_nfock ? l2g_gen_fock[c] : l2g_gen_fock32[c - l2g_gen_nfock];
}
int fock_ybuf(int c) { return fock_is_generic(c) ? 0 : fock_info(c).ybuf_doubles; }

// The screened unique quartets of one Fock build, by kernel class.
struct FockWork {
  std::vector<FockItems> cls;
  int64_t nq = 0, nq32 = 0, nprim = 0;
};

// Reorders the items of one class by decreasing primitive-loop length, so
// that neighbouring lanes run equal-length loops.  Stable counting sort.
void sort_by_depth(FockItems& it) {
  const size_t m = it.size();
  if (m < 2) return;
  int wmax = 0;
  std::vector<int> w(m);
  for (size_t i = 0; i < m; ++i) {
    w[i] = it.pr[i].y - it.pr[i].x;
    wmax = std::max(wmax, w[i]);
  }
  std::vector<size_t> cnt((size_t)wmax + 2, 0);
  for (size_t i = 0; i < m; ++i) ++cnt[(size_t)(wmax - w[i]) + 1];
  for (size_t k = 1; k < cnt.size(); ++k) cnt[k] += cnt[k - 1];
  std::vector<size_t> ord(m);
  for (size_t i = 0; i < m; ++i) ord[cnt[(size_t)(wmax - w[i])]++] = i;
  FockItems s;
  s.bra.resize(m), s.ket.resize(m), s.meta.resize(m), s.lnp.resize(m), s.scale.resize(m), s.bf.resize(m),
      s.pr.resize(m);
  if (!it.req.empty()) s.req.resize(m);
  for (size_t k = 0; k < m; ++k) {
    const size_t i = ord[k];
    s.bra[k] = it.bra[i], s.ket[k] = it.ket[i], s.meta[k] = it.meta[i];
    s.lnp[k] = it.lnp[i], s.scale[k] = it.scale[i], s.bf[k] = it.bf[i], s.pr[k] = it.pr[i];
    if (!it.req.empty()) s.req[k] = it.req[i];
  }
  it = std::move(s);
}

// Classes with fewer items than the chip has lanes to fill are split by
// primitive-quartet range: an item whose loop is longer than the class's
// chunk (about 1.4e5 recursion-node evaluations) becomes several items on
// several lanes.  The digest is linear in the integrals, so each range is
// digested on its own; the cost is one more HRR, transform and digest per
// range.  Without this, one (dd|dd) quartet with K = 6 on every shell runs
// ~1300 primitive quartets x 19 groups on one lane, tens of ms.
// Longest primitive range of one work item of class cls, for m items with
// `total` primitive quartets in all, on a chip of `lanes` lanes (one wave
// per SIMD).  A lane runs its primitive loop serially, each iteration waiting
// on table and pair loads, so a deep item would bound the whole class: items
// are cut to half of the class's per-lane share, but not below 8-16
// primitives (fewer for classes with large recursions), which keeps the
// per-piece HRR and digest small against the loop.  The generic kernel runs
// one item per block; only its classes with few items are split, to about
// 2e6 recursion values per piece.
int split_chunk(int cls, int64_t m, int64_t total, int lanes) {
  if (m == 0) return INT_MAX;
  if (fock_is_generic(cls))
    return m < 4096 ? std::max(16, 2000000 / std::max(1, fock_gclass(cls).nvals)) : INT_MAX;
  const int minc = std::max(8, std::min(16, 140000 / std::max(1, fock_info(cls).nodes)));
  const int64_t share = (total + 2ll * lanes - 1) / (2ll * std::max(1, lanes));
  return (int)std::min<int64_t>(INT_MAX, std::max<int64_t>(minc, share));
}

void split_deep(FockItems& it, int cls, int lanes) {
  const size_t m = it.size();
  int64_t total = 0;
  for (size_t i = 0; i < m; ++i) total += it.pr[i].y - it.pr[i].x;
  const int chunk = split_chunk(cls, (int64_t)m, total, lanes);
  if (chunk == INT_MAX) return;
  size_t extra = 0;
  for (size_t i = 0; i < m; ++i) {
    const int len = it.pr[i].y - it.pr[i].x;
    if (len > chunk) extra += (size_t)((len + chunk - 1) / chunk) - 1;
  }
  if (extra == 0) return;
  FockItems s;
  s.bra.reserve(m + extra), s.ket.reserve(m + extra), s.meta.reserve(m + extra), s.lnp.reserve(m + extra),
      s.scale.reserve(m + extra), s.bf.reserve(m + extra), s.pr.reserve(m + extra);
  if (!it.req.empty()) s.req.reserve(m + extra);
  for (size_t i = 0; i < m; ++i) {
    const int b = it.pr[i].x, e = it.pr[i].y, len = e - b;
    const int nch = len > chunk ? (len + chunk - 1) / chunk : 1;
    for (int k = 0; k < nch; ++k) {
      s.bra.push_back(it.bra[i]), s.ket.push_back(it.ket[i]), s.meta.push_back(it.meta[i]);
      s.lnp.push_back(it.lnp[i]), s.scale.push_back(it.scale[i]), s.bf.push_back(it.bf[i]);
      if (!it.req.empty()) s.req.push_back(it.req[i]);
      s.pr.push_back(int2{b + (int)((int64_t)len * k / nch), b + (int)((int64_t)len * (k + 1) / nch)});
    }
  }
  it = std::move(s);
}

// Screened, permutationally-unique shell quartets of compute_2body_fock
// (hartree-fock++.cc:1628-1696) for density D, bucketed by kernel class:
// shell-block norms of D, the Schwarz test Dnorm1234 K12 K34 >= precision,
// per-quartet engine precision precision / Dnorm1234 and the degeneracy
// factor; canonical shell order as engine.impl.h:1181-1208 with equal-l
// pairs oriented by second_is_centre.  FP32 classes take the quartets with
// K12 K34 < fp32_threshold and S12 S34 < L2G_FP32_HRR_FACTOR fp32_threshold.
// Runs on all host threads over blocks of s1; the block results are joined
// in s1 order, so the item order does not depend on the thread count.
void fock_enumerate(const FockSetup& f, const double* D, double precision, double fp32_threshold, FockWork& out) {
  l2g_basis_s* B = f.B;
  l2g_context_s* ctx = B->ctx;
  const int n = f.n, N = f.nbf;
  std::vector<double> Dn((size_t)n * n, 0.0);
  parallel_for(n, [&](int s1) {
    for (int s2 = 0; s2 < n; ++s2) {
      double m = 0;
      for (int i = f.bf1st[s1]; i < f.bf1st[s1 + 1]; ++i)
        for (int j = f.bf1st[s2]; j < f.bf1st[s2 + 1]; ++j) m = std::max(m, std::fabs(D[(size_t)i * N + j]));
      Dn[(size_t)s1 * n + s2] = m;
    }
  });
  double dmax = 0;
  for (double v : Dn) dmax = std::max(dmax, v);
  const double needed = dmax > 0 ? precision / dmax : precision;
  if (needed < f.max_engine_precision)
    throw std::invalid_argument("requested Fock precision needs more engine precision than the shell-pair data hold");
  const int ncls = nfock_classes();
  const int nblk = std::max(1, std::min(n, 512));
  std::vector<FockWork> part(nblk);
  parallel_for(nblk, [&](int blk) {
    FockWork& w = part[blk];
    w.cls.assign(ncls, {});
    const int lo = (int)((int64_t)n * blk / nblk), hi = (int)((int64_t)n * (blk + 1) / nblk);
    for (int s1 = lo; s1 < hi; ++s1)
      for (int s2 : f.splist[s1]) {
        const double Dn12 = Dn[(size_t)s1 * n + s2];
        const double K12 = f.K[(size_t)s1 * n + s2], S12 = f.S[(size_t)s1 * n + s2];
        for (int s3 = 0; s3 <= s1; ++s3) {
          const double Dn123 = std::max(std::max(Dn[(size_t)s1 * n + s3], Dn[(size_t)s2 * n + s3]), Dn12);
          const int s4max = (s1 == s3) ? s2 : s3;
          for (int s4 : f.splist[s3]) {
            if (s4 > s4max) break;
            const double Dn1234 = std::max(std::max(Dn[(size_t)s1 * n + s4], Dn[(size_t)s2 * n + s4]),
                                           std::max(Dn[(size_t)s3 * n + s4], Dn123));
            const double K34 = f.K[(size_t)s3 * n + s4];
            if (Dn1234 * K12 * K34 < precision) continue;
            const double qprec = Dn1234 != 0. ? precision / Dn1234 : needed;
            const int deg = (s1 == s2 ? 1 : 2) * (s3 == s4 ? 1 : 2) * (s1 == s3 ? (s2 == s4 ? 1 : 2) : 2);
            const int l1 = B->l[s1], l2 = B->l[s2], l3 = B->l[s3], l4 = B->l[s4];
            const bool stb = B->swap_pair(s1, s2), stk = B->swap_pair(s3, s4), sbk = (l1 + l2) > (l3 + l4);
            const int b1 = sbk ? (stk ? s4 : s3) : (stb ? s2 : s1);
            const int b2 = sbk ? (stk ? s3 : s4) : (stb ? s1 : s2);
            const int k1 = sbk ? (stb ? s2 : s1) : (stk ? s4 : s3);
            const int k2 = sbk ? (stb ? s1 : s2) : (stk ? s3 : s4);
            int cls = ctx->force_generic ? -1 : ctx->fock_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
            if (cls < 0) {
              const int g = ctx->generic_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
              if (g < 0) throw L2gError(L2G_E_LMAX, "no Fock kernel for this angular momentum class");
              cls = l2g_gen_nfock + l2g_gen_nfock32 + g;
            } else if (fp32_threshold > 0 && K12 * K34 < fp32_threshold &&
                S12 * f.S[(size_t)s3 * n + s4] < L2G_FP32_HRR_FACTOR * fp32_threshold) {
              const int c32 = ctx->fock32_of[B->l[b1]][B->l[b2]][B->l[k1]][B->l[k2]];
              if (c32 >= 0 && ctx->fp32_policy) {
                cls = l2g_gen_nfock + c32;
                ++w.nq32;
              }
            }
            auto purebit = [&](int sh) { return (B->pure[sh] && B->l[sh] > 0) ? 1 : 0; };
            FockItems& it = w.cls[cls];
            const int bp = f.pair_of[(size_t)b1 * n + b2], kp = f.pair_of[(size_t)k1 * n + k2];
            it.bra.push_back(bp);
            it.ket.push_back(kp);
            it.meta.push_back((purebit(b1) << 3) | (purebit(b2) << 4) | (purebit(k1) << 5) | (purebit(k2) << 6));
            it.lnp.push_back(qprec > 0 ? std::log(qprec) : std::numeric_limits<double>::lowest());
            it.scale.push_back((double)deg);
            it.bf.push_back(int4{f.bf1st[b1], f.bf1st[b2], f.bf1st[k1], f.bf1st[k2]});
            it.pr.push_back(int2{0, f.np[bp] * f.np[kp]});
            ++w.nq;
            w.nprim += (int64_t)f.np[bp] * f.np[kp];
          }
        }
      }
  });
  out.cls.assign(ncls, {});
  out.nq = out.nq32 = out.nprim = 0;
  for (auto& w : part) {
    out.nq += w.nq;
    out.nq32 += w.nq32;
    out.nprim += w.nprim;
  }
  parallel_for(ncls, [&](int c) {
    size_t tot = 0;
    for (auto& w : part) tot += w.cls[c].size();
    if (tot == 0) return;
    FockItems& o = out.cls[c];
    o.bra.reserve(tot), o.ket.reserve(tot), o.meta.reserve(tot), o.lnp.reserve(tot), o.scale.reserve(tot),
        o.bf.reserve(tot);
    for (auto& w : part) o.append(w.cls[c]);
    split_deep(o, c, f.B->ctx->ncu * 256);
    sort_by_depth(o);
  });
}

// Device copies of a FockSetup's shell-pair data.
struct PairDev {
  DevBuf first, np, AB, P, PA, aux, ns;
  void upload_from(const FockSetup& f, hipStream_t st) {
    upload(first, f.first, st);
    upload(np, f.np, st);
    upload(AB, f.AB, st);
    upload(P, f.P, st);
    upload(PA, f.PA, st);
    upload(aux, f.aux, st);
    upload(ns, f.ns, st);
  }
};

// Kernel arguments common to every class of one Fock launch round.
L2gEriArgs fock_args(l2g_context_s* ctx, int method, const int* pair_first, const int* pair_np, const double* pair_AB,
                     const double4* P, const double4* PA, const double2* aux, const double* ns, double* dens,
                     double* jg, double* kg, const L2gMatDesc* mats, double* ybuf, int64_t ystride,
                     long long fx_hoff) {
  L2gEriArgs A;
  std::memset(&A, 0, sizeof(A));
  A.pair_first = pair_first;
  A.pair_np = pair_np;
  A.pair_AB = pair_AB;
  A.pp_P = P;
  A.pp_PA = PA;
  A.pp_aux = aux;
  A.pp_nonsph = ns;
  A.boys = ctx->boys.as<double>();
  A.scale = 1.0;
  A.method = method;
  A.shg_gaussian = ctx->shg_gaussian;
  A.dens = dens;
  A.jg = jg;
  A.kg = kg;
  A.mats = mats;
  A.ybuf = ybuf;
  A.ystride = ystride;
  A.fx_hoff = fx_hoff;
  return A;
}

// Concatenated device copy of the work items of all classes, with the class
// ranges; the arrays are uploaded from host vectors that must stay alive
// until the stream has passed the copies.
// Device work items of one launch round: arrays and per-class ranges.
struct ItemsView {
  const int *bra = nullptr, *ket = nullptr, *meta = nullptr, *req = nullptr;
  const double *lnp = nullptr, *scale = nullptr;
  const int4* bf = nullptr;
  const int2* pr = nullptr;
  std::vector<int64_t> first, count;  // per class
};

struct ItemsDev {
  DevBuf bra, ket, meta, req, lnp, scale, bf, pr;
  std::vector<int64_t> first, count;  // per class
  FockItems host;                      // staging (kept until the stream is synchronised)
  ItemsView view() const {
    ItemsView v;
    v.bra = bra.as<int>(), v.ket = ket.as<int>(), v.meta = meta.as<int>(), v.req = req.as<int>();
    v.lnp = lnp.as<double>(), v.scale = scale.as<double>(), v.bf = bf.as<int4>(), v.pr = pr.as<int2>();
    v.first = first, v.count = count;
    return v;
  }
  void upload_from(const std::vector<FockItems>& cls, hipStream_t st) {
    host.clear();
    first.assign(cls.size(), 0);
    count.assign(cls.size(), 0);
    for (size_t c = 0; c < cls.size(); ++c) {
      first[c] = (int64_t)host.size();
      count[c] = (int64_t)cls[c].size();
      host.append(cls[c]);
    }
    upload(bra, host.bra, st);
    upload(ket, host.ket, st);
    upload(meta, host.meta, st);
    upload(lnp, host.lnp, st);
    upload(scale, host.scale, st);
    upload(bf, host.bf, st);
    upload(pr, host.pr, st);
    if (!host.req.empty()) upload(req, host.req, st);
  }
};

// Launches every non-empty class of `items` with the common arguments A;
// grid and scratch as sized by the caller.
// With `ev` (2 events per class), records the class's launch between
// ev[2c] and ev[2c+1].
// Launches every class of an item set.  With slots (plan_slots over the
// classes), each class gets its own grid and scratch region, and the classes
// too small to fill the chip run on the side streams beside each other;
// without, every class runs in order on st with the scratch in A.
void fock_launch_all(l2g_context_s* ctx, L2gEriArgs A, const ItemsView& items, bool with_req, int grid, int block,
                     hipStream_t st, hipEvent_t* ev = nullptr, const std::vector<ClassSlot>* slots = nullptr) {
  double* const ybase = A.ybuf;
  bool sides = false;
  if (slots)
    for (size_t c = 0; c < items.count.size(); ++c) sides |= items.count[c] > 0 && (*slots)[c].side >= 0;
  if (sides) fork_sides(*ctx);
  try {
  for (size_t c = 0; c < items.count.size(); ++c) {
    const int64_t m = items.count[c], o = items.first[c];
    if (m == 0) continue;
    hipStream_t cs = st;
    int g = (int)std::min<int64_t>(grid, (m + block - 1) / block);
    if (slots) {
      const ClassSlot& sl = (*slots)[c];
      if (sides && sl.side >= 0) cs = ctx->side[sl.side];
      A.ybuf = ybase + sl.yoff;
      A.ystride = sl.ystride;
      g = sl.grid;
    }
    if (ev) HIPCHK(hipEventRecord(ev[2 * c], cs));
    A.bra_pair = items.bra + o;
    A.ket_pair = items.ket + o;
    A.meta = items.meta + o;
    A.q_lnprec = items.lnp + o;
    A.q_scale = items.scale + o;
    A.q_bf = items.bf + o;
    A.q_req = with_req ? items.req + o : nullptr;
    A.q_prange = items.pr + o;
    A.n = (int)m;
    if (fock_is_generic((int)c)) {
      launch_generic(kGenericFock, A, fock_gclass((int)c), ctx->gprog.as<int4>(), ctx->gidx.as<int>(),
                     (int)std::min<int64_t>((int64_t)ctx->ncu * 8, m), cs);
    } else {
      fock_info((int)c).launch(A, g, block, cs);
      LAUNCHCHK(std::string("Fock kernel of class ") + std::to_string(c) + " (grid " + std::to_string(g) +
                ", block " + std::to_string(block) + ", items " + std::to_string(m) + ")");
    }
    if (ev) HIPCHK(hipEventRecord(ev[2 * c + 1], cs));
  }
  } catch (...) {
    // the side streams are joined back also when a launch fails
    if (sides) {
      try {
        join_sides(*ctx);
      } catch (...) {
      }
    }
    throw;
  }
  if (sides) join_sides(*ctx);
}

}  // namespace

// quartets per window of a Fock build's list (about 200 bytes of device
// memory each while the window is sorted and split)
#ifndef L2G_FOCK_WINDOW_ITEMS
#define L2G_FOCK_WINDOW_ITEMS (1ll << 26)
#endif

struct l2g_fock_s {
  FockSetup setup;
  long long window_items = L2G_FOCK_WINDOW_ITEMS;
  double fp32_threshold = 0;  // FP32 selection threshold (fock_enumerate)
  bool deterministic = false;  // fixed-point J/K accumulation (l2g_fx_add)
  l2g_progress_fn progress = nullptr;  // called after each window
  void* progress_user = nullptr;
  int64_t last_nq = 0, last_nprim = 0, last_nq32 = 0;
  int64_t last_windows = 0;  // windows the last build's quartet list came in
  float last_ms = 0;
  std::vector<float> cls_ms;  // per class: kernel ms of the last build, over its windows
  PairDev pairs;  // resident shell-pair data
  DevBuf dens, jg, kg, mats, ybuf;
  // GPU quartet enumeration: pair list in (s1, s2) order, per-shell data,
  // class maps, shell-block norms of D and the sorted items
  DevBuf e_ps1, e_ps2, e_pK, e_pS, e_pid, e_pswap, e_l, e_pure, e_bf1st, e_np, e_cls, e_cls32, e_Dn;
  int64_t npairs = 0;
  FockEnumState enumst;
  std::vector<hipEvent_t> ev;     // per class: start, end of the last build
  std::vector<int64_t> cls_nq, cls_nprim;
  ~l2g_fock_s() {
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

// Device tables of the GPU quartet enumeration (l2g_enum.hip): the
// significant pairs in (s1, s2) order with their Schwarz and FP32-selection
// factors, canonical pair ids and orientation, per-shell data, and the map
// from a canonical class to its Fock kernel class.
void upload_enum_tables(l2g_fock_s& f, hipStream_t st) {
  const FockSetup& fs = f.setup;
  l2g_basis_s* B = fs.B;
  l2g_context_s* ctx = B->ctx;
  const int n = fs.n;
  std::vector<int> ps1, ps2, pid;
  std::vector<double> pK, pS;
  std::vector<unsigned char> pswap;
  for (int s1 = 0; s1 < n; ++s1)
    for (int s2 : fs.splist[s1]) {
      ps1.push_back(s1);
      ps2.push_back(s2);
      pK.push_back(fs.K[(size_t)s1 * n + s2]);
      pS.push_back(fs.S[(size_t)s1 * n + s2]);
      pid.push_back(fs.pair_of[(size_t)s1 * n + s2]);
      pswap.push_back(B->swap_pair(s1, s2) ? 1 : 0);
    }
  f.npairs = (int64_t)ps1.size();
  std::vector<int> cls(256, -1), cls32(256, -1);
  const int nf64 = l2g_gen_nfock, nf32 = l2g_gen_nfock32;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) {
          const int k = ((a * 4 + b) * 4 + c) * 4 + d;
          const int spec = ctx->force_generic ? -1 : ctx->fock_of[a][b][c][d];
          const int g = ctx->generic_of[a][b][c][d];
          cls[k] = spec >= 0 ? spec : (g >= 0 ? nf64 + nf32 + g : -1);
          const int c32 = ctx->force_generic ? -1 : ctx->fock32_of[a][b][c][d];
          // FP32 Fock kernels only when asked for every class: in mixed Fock
          // builds the FP32-eligible quartets are the light ones (few
          // primitive quartets survive), where FP32 saves nothing and the
          // extra launches cost (measured 0.88-0.95x)
          cls32[k] = c32 >= 0 && ctx->fp32_policy ? nf64 + c32 : -1;
        }
  if (nfock_classes() > kEnumClasses - 1) throw L2gError(L2G_E_STATE, "too many Fock kernel classes for the enumeration sort key");
  upload(f.e_ps1, ps1, st);
  upload(f.e_ps2, ps2, st);
  upload(f.e_pK, pK, st);
  upload(f.e_pS, pS, st);
  upload(f.e_pid, pid, st);
  upload(f.e_pswap, pswap, st);
  upload(f.e_l, B->l, st);
  upload(f.e_pure, B->pure, st);
  upload(f.e_bf1st, fs.bf1st, st);
  upload(f.e_np, fs.np, st);
  upload(f.e_cls, cls, st);
  upload(f.e_cls32, cls32, st);
  HIPCHK(hipStreamSynchronize(st));  // the host vectors go out of scope
}

}  // namespace

extern "C" {

int l2g_fock_create(l2g_basis B, double pair_threshold, double max_engine_precision, int screening, l2g_fock* out) {
  return guarded([&] {
    if (!B || !out) throw std::invalid_argument("l2g_fock_create: null argument");
    HIPCHK(hipSetDevice(B->ctx->device));
    auto f = std::make_unique<l2g_fock_s>();
    fock_setup(f->setup, B, pair_threshold, max_engine_precision, screening);
    f->pairs.upload_from(f->setup, B->ctx->stream);
    upload_enum_tables(*f, B->ctx->stream);
    const L2gMatDesc md{0, f->setup.nbf};
    f->mats.alloc(sizeof(md));
    HIPCHK(hipMemcpyAsync(f->mats.p, &md, sizeof(md), hipMemcpyHostToDevice, B->ctx->stream));
    HIPCHK(hipStreamSynchronize(B->ctx->stream));
    ++B->ctx->live;
    basis_retain(B);
    *out = f.release();
  });
}

int l2g_fock_nbf(l2g_fock f, int* nbf) {
  return guarded([&] {
    if (!f || !nbf) throw std::invalid_argument("null argument");
    *nbf = f->setup.nbf;
  });
}

int l2g_fock_schwarz(l2g_fock f, double* K) {
  return guarded([&] {
    if (!f || !K) throw std::invalid_argument("null argument");
    std::memcpy(K, f->setup.K.data(), f->setup.K.size() * sizeof(double));
  });
}

}  // extern "C"

namespace {

// The count pass of a build: shell-block norms of D, then the screened
// unique quartets counted on the GPU per enumeration block (f->enumst).
FockEnumArgs fock_count_pass(l2g_fock_s* f, const double* D, double precision) {
  const FockSetup& fs = f->setup;
  l2g_context_s* ctx = fs.B->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  const int n = fs.n, N = fs.nbf;
  hipStream_t st = ctx->stream;
  // shell-block infinity norms of D (compute_shellblock_norm)
  std::vector<double> Dn((size_t)n * n, 0.0);
  parallel_for(n, [&](int s1) {
    for (int s2 = 0; s2 < n; ++s2) {
      double m = 0;
      for (int i = fs.bf1st[s1]; i < fs.bf1st[s1 + 1]; ++i)
        for (int j = fs.bf1st[s2]; j < fs.bf1st[s2 + 1]; ++j) m = std::max(m, std::fabs(D[(size_t)i * N + j]));
      Dn[(size_t)s1 * n + s2] = m;
    }
  });
  double dmax = 0;
  for (double v : Dn) dmax = std::max(dmax, v);
  const double needed = dmax > 0 ? precision / dmax : precision;
  if (needed < fs.max_engine_precision)
    throw std::invalid_argument("requested Fock precision needs more engine precision than the shell-pair data hold");
  upload(f->e_Dn, Dn, st);
  // screened unique quartets on the GPU, sorted by class and depth
  FockEnumArgs E;
  E.ps1 = f->e_ps1.as<int>();
  E.ps2 = f->e_ps2.as<int>();
  E.pK = f->e_pK.as<double>();
  E.pS = f->e_pS.as<double>();
  E.pid = f->e_pid.as<int>();
  E.pswap = f->e_pswap.as<unsigned char>();
  E.l = f->e_l.as<int>();
  E.pure = f->e_pure.as<int>();
  E.bf1st = f->e_bf1st.as<int>();
  E.np = f->e_np.as<int>();
  E.Dn = f->e_Dn.as<double>();
  E.n = n;
  E.npairs = f->npairs;
  E.clsmap = f->e_cls.as<int>();
  E.cls32map = f->e_cls32.as<int>();
  E.precision = precision;
  E.needed = needed;
  E.fp32_thr = f->fp32_threshold;
  E.hrr_factor = L2G_FP32_HRR_FACTOR;
  FockEnumState& es = f->enumst;
  fock_count_device(E, es, st);  // synchronises; Dn may go
  return E;
}

}  // namespace

extern "C" {

int l2g_fock_count(l2g_fock f, const double* D, double precision, int64_t* nquartets) {
  return guarded([&] {
    if (!f || !D || !nquartets) throw std::invalid_argument("l2g_fock_count: null argument");
    (void)fock_count_pass(f, D, precision);
    long long n = 0;
    for (long long c : f->enumst.block_count) n += c;
    *nquartets = n;
  });
}

int l2g_fock_build(l2g_fock f, const double* D, double precision, double* J, double* K, double* G) {
  return guarded([&] {
    if (!f || !D) throw std::invalid_argument("l2g_fock_build: null argument");
    const FockSetup& fs = f->setup;
    l2g_context_s* ctx = fs.B->ctx;
    const int N = fs.nbf;
    hipStream_t st = ctx->stream;
    FockEnumArgs E = fock_count_pass(f, D, precision);
    FockEnumState& es = f->enumst;
    long long total_nq = 0;
    for (long long c : es.block_count) total_nq += c;
    const int ncls = nfock_classes();
    f->last_nq = 0;
    f->last_nq32 = 0;
    f->last_nprim = 0;
    f->last_ms = 0;
    f->last_windows = 0;
    f->cls_nq.assign(ncls, 0);
    f->cls_nprim.assign(ncls, 0);
    f->cls_ms.assign(ncls, 0.0f);
    const size_t NN = (size_t)N * N;
    f->dens.alloc(std::max<size_t>(NN, 1) * sizeof(double));
    // deterministic digest: a plane of low words, then one of high words
    const size_t planes = f->deterministic ? 2 : 1;
    const long long hoff = f->deterministic ? (long long)NN : 0;
    f->jg.alloc(std::max<size_t>(planes * NN, 1) * sizeof(double));
    f->kg.alloc(std::max<size_t>(planes * NN, 1) * sizeof(double));
    if (NN) {
      HIPCHK(hipMemcpyAsync(f->dens.p, D, NN * sizeof(double), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemsetAsync(f->jg.p, 0, planes * NN * sizeof(double), st));
      HIPCHK(hipMemsetAsync(f->kg.p, 0, planes * NN * sizeof(double), st));
    }
    if (f->ev.empty()) {
      f->ev.assign(2 * ncls + 2, nullptr);
      for (auto& e : f->ev) HIPCHK(hipEventCreate(&e));
    }
    const int block = 128;
    const int grid = (int)ctx->ncu * 8;
    // The quartet list is produced and digested in windows of consecutive
    // count-pass blocks holding at most window_items quartets each (one
    // block may exceed it alone), so that a large molecule's list (1e10
    // quartets and more) never has to sit in device memory at once.
    int blk = 0;
    while (blk < es.nblocks) {
      int blk1 = blk + 1;
      long long items = es.block_count[blk];
      while (blk1 < es.nblocks && items + es.block_count[blk1] <= f->window_items) items += es.block_count[blk1++];
      fock_fill_device(E, es, blk, blk1, st);
      blk = blk1;
      if (es.n == 0) continue;
      ++f->last_windows;
      f->last_nq += es.n;
      std::vector<int> chunk(kEnumClasses, INT_MAX);
      for (int c = 0; c < ncls; ++c) {
        const int64_t m = (int64_t)es.cls_count[c];
        f->cls_nq[c] += m;
        f->cls_nprim[c] += (int64_t)es.cls_nprim[c];
        f->last_nprim += (int64_t)es.cls_nprim[c];
        if (c >= l2g_gen_nfock && !fock_is_generic(c)) f->last_nq32 += m;
        chunk[c] = split_chunk(c, m, (int64_t)es.cls_nprim[c], ctx->ncu * 256);
      }
      // deep items cut by primitive range on the device (class-major order kept)
      fock_split_device(es, chunk, st);
      ItemsView big;
      big.bra = es.bra.as<int>(), big.ket = es.ket.as<int>(), big.meta = es.meta.as<int>();
      big.lnp = es.lnp.as<double>(), big.scale = es.scale.as<double>(), big.bf = es.bf.as<int4>();
      big.pr = es.pr.as<int2>();
      big.first.assign(ncls, 0);
      big.count.assign(ncls, 0);
      int64_t o = 0;
      for (int c = 0; c < ncls; ++c) {
        const int64_t m = (int64_t)es.cls_pieces[c];
        big.first[c] = o;
        big.count[c] = m;
        o += m;
      }
      int64_t ymax = 0;
      std::vector<int64_t> lanes(ncls, 0), yd(ncls, 0);
      for (int c = 0; c < ncls; ++c) {
        const int64_t m = big.count[c];
        if (m == 0) continue;
        // the generic kernel runs a block per item, no lane scratch, a few
        // blocks per CU: up to 4 items per CU count as small (side streams)
        lanes[c] = fock_is_generic(c) ? m * 64 : m;
        yd[c] = fock_ybuf(c);
        ymax = std::max<int64_t>(ymax, yd[c]);
      }
      std::vector<ClassSlot> slots;
      const int64_t ytotal = plan_slots(slots, lanes, yd, ctx->ncu, block, l2g_context_s::kSide,
                                        ymax * (int64_t)ctx->ncu * 8 * block + (int64_t(1) << 29));
      f->ybuf.alloc(std::max<int64_t>(ytotal, 1) * sizeof(double));
      const int64_t ystride = 0;  // per class, from its slot
      const PairDev& p = f->pairs;
      L2gEriArgs A = fock_args(ctx, fs.method, p.first.as<int>(), p.np.as<int>(), p.AB.as<double>(),
                               p.P.as<double4>(), p.PA.as<double4>(), p.aux.as<double2>(), p.ns.as<double>(),
                               f->dens.as<double>(), f->jg.as<double>(), f->kg.as<double>(), f->mats.as<L2gMatDesc>(),
                               f->ybuf.as<double>(), ystride, hoff);
      hipEvent_t e0 = f->ev[2 * ncls], e1 = f->ev[2 * ncls + 1];
      HIPCHK(hipEventRecord(e0, st));
      fock_launch_all(ctx, A, big, false, grid, block, st, f->ev.data(), &slots);
      HIPCHK(hipEventRecord(e1, st));
      // the window's items and scratch are reused by the next window
      hipError_t e = hipStreamSynchronize(st);
      if (e != hipSuccess) throw L2gError(L2G_E_LAUNCH, std::string("Fock kernels failed: ") + hipGetErrorString(e));
      f->last_ms += event_ms_or(e0, e1, 0.0f);
      for (int c = 0; c < ncls; ++c)
        if (big.count[c]) f->cls_ms[c] += event_ms_or(f->ev[2 * c], f->ev[2 * c + 1], 0.0f);
      if (f->progress) f->progress(f->last_nq, total_nq, f->progress_user);
    }
    launch_fock_finish(f->jg.as<double>(), f->kg.as<double>(), f->dens.as<double>(), N, hoff, st);
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) throw L2gError(L2G_E_LAUNCH, std::string("Fock kernels failed: ") + hipGetErrorString(e));
    if (J && NN) HIPCHK(hipMemcpy(J, f->jg.p, NN * sizeof(double), hipMemcpyDeviceToHost));
    if (K && NN) HIPCHK(hipMemcpy(K, f->kg.p, NN * sizeof(double), hipMemcpyDeviceToHost));
    if (G && NN) HIPCHK(hipMemcpy(G, f->dens.p, NN * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int l2g_fock_set_progress(l2g_fock f, l2g_progress_fn fn, void* user) {
  return guarded([&] {
    if (!f) throw std::invalid_argument("l2g_fock_set_progress: null Fock builder");
    f->progress = fn;
    f->progress_user = user;
  });
}

int l2g_fock_set_deterministic(l2g_fock f, int on) {
  return guarded([&] {
    if (!f) throw std::invalid_argument("l2g_fock_set_deterministic: null Fock builder");
    f->deterministic = on != 0;
  });
}

int l2g_fock_set_window(l2g_fock f, int64_t items) {
  return guarded([&] {
    if (!f || items < 0) throw std::invalid_argument("l2g_fock_set_window: bad arguments");
    f->window_items = items ? items : L2G_FOCK_WINDOW_ITEMS;
  });
}

int l2g_fock_stats_windows(l2g_fock f, int64_t* windows) {
  return guarded([&] {
    if (!f || !windows) throw std::invalid_argument("null argument");
    *windows = f->last_windows;
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

int l2g_fock_class_stats(l2g_fock f, int cap, int* nclasses, int* class_code, int64_t* nquartets,
                         int64_t* nprim_quartets, float* ms) {
  return guarded([&] {
    if (!f || !nclasses || cap < 0) throw std::invalid_argument("l2g_fock_class_stats: bad arguments");
    int k = 0;
    for (int c = 0; c < (int)f->cls_nq.size(); ++c) {
      if (f->cls_nq[c] == 0) continue;
      if (k < cap) {
        int la, lb, lc, ld;
        if (fock_is_generic(c)) {
          const L2gGenericClass& g = fock_gclass(c);
          la = g.la, lb = g.lb, lc = g.lc, ld = g.ld;
        } else {
          const L2gClassInfo& ci = fock_info(c);
          la = ci.la, lb = ci.lb, lc = ci.lc, ld = ci.ld;
        }
        if (class_code)
          class_code[k] = la * 1000 + lb * 100 + lc * 10 + ld +
                          (fock_is_generic(c) ? 20000 : (c >= l2g_gen_nfock ? 10000 : 0));
        if (nquartets) nquartets[k] = f->cls_nq[c];
        if (nprim_quartets) nprim_quartets[k] = f->cls_nprim[c];
        if (ms) ms[k] = c < (int)f->cls_ms.size() ? f->cls_ms[c] : 0.0f;
      }
      ++k;
    }
    *nclasses = k;
  });
}

int l2g_fock_destroy(l2g_fock f) {
  if (f && !g_exiting) {
    l2g_basis_s* B = f->setup.B;
    l2g_context_s* ctx = B->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete f;
    --ctx->live;
    basis_release(B);  // the builder's hold on its basis
  }
  return L2G_OK;
}

}  // extern "C"

// ===========================================================================
// Fock-build request server
// ===========================================================================
namespace {

// First-fit allocator of element ranges in one device arena.
struct FreeList {
  std::map<int64_t, int64_t> blocks;  // offset -> length of each free block
  void init(int64_t cap) {
    blocks.clear();
    if (cap > 0) blocks[0] = cap;
  }
  // offset of a free range of len elements, or -1
  int64_t alloc(int64_t len) {
    if (len <= 0) return 0;
    for (auto it = blocks.begin(); it != blocks.end(); ++it)
      if (it->second >= len) {
        const int64_t off = it->first, rest = it->second - len;
        blocks.erase(it);
        if (rest > 0) blocks[off + len] = rest;
        return off;
      }
    return -1;
  }
  void release(int64_t off, int64_t len) {
    if (len <= 0 || off < 0) return;
    auto it = blocks.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != blocks.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      blocks.erase(nx);
    }
    if (it != blocks.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        blocks.erase(it);
      }
    }
  }
};

using Clock = std::chrono::steady_clock;

struct ServerRequest {
  int64_t id = 0;
  int state = L2G_REQ_QUEUED;
  bool claimed = false;  // an admission worker is preparing it
  int error = 0;
  std::string message;
  // inputs (copied at submission)
  std::vector<int> l, pure, nprim;
  std::vector<double> alpha, coeff, origin, D;
  int nbf = 0;
  double precision = 0, fp32 = 0;
  // admission
  l2g_basis basis = nullptr;
  FockSetup setup;
  FockWork work;
  int slot = -1;
  int64_t pair_off = -1, prim_off = -1, mat_off = -1;
  int64_t npair = 0, nprimpair = 0, nmat = 0;
  // launch progress (launch thread only)
  std::vector<size_t> cursor;
  int64_t left = 0;  // work items not yet launched
  hipEvent_t done_ev = nullptr;
  double* result = nullptr;  // pinned host: J, K, G
  Clock::time_point t_submit, t_admit, t_done;
};

}  // namespace

struct l2g_server_s {
  l2g_server_config cfg;
  l2g_context exec = nullptr;   // launch rounds
  l2g_context admit = nullptr;  // Schwarz kernels and uploads of admissions
  // device arenas shared by all admitted requests
  DevBuf a_first, a_np, a_AB, a_P, a_PA, a_aux, a_ns, a_dens, a_jg, a_kg, a_mats, ybuf;
  long long fx_hoff = 0;  // deterministic digest: words from a low word to its high word (0: FP64 atomics)
  FreeList fl_pair, fl_prim, fl_mat;
  int64_t ystride = 0;
  int grid = 1;
  const int block = 128;
  std::vector<int> free_slots;
  // launch-round buffers, double-buffered: a buffer is refilled only after
  // the event of the round that last used it has completed
  ItemsDev rbuf[2];
  std::vector<L2gMatDesc> rmats_host[2];
  DevBuf rmats[2];
  hipEvent_t rev[2] = {nullptr, nullptr};
  int rnext = 0;
  // requests (guarded by mu)
  std::mutex mu;
  std::condition_variable cv;
  std::map<int64_t, std::unique_ptr<ServerRequest>> reqs;
  std::deque<ServerRequest*> pending;      // submitted, not admitted
  std::vector<ServerRequest*> active;      // admitted, quartets left to launch
  std::vector<ServerRequest*> finishing;   // all launched, finish step queued
  int64_t next_id = 1, inflight = 0;
  int uploading = 0;  // reserved requests between reservation and activation
  bool stop = false;
  std::vector<std::thread> admit_thr;  // admission workers
  std::thread exec_thr;
  int64_t rounds = 0, max_per_round = 0, nq = 0;
};

namespace {

// live servers, so that the exit handler can stop their threads
std::mutex g_servers_mu;
std::vector<l2g_server_s*> g_servers;

void server_registry(l2g_server_s* S, bool add) {
  std::lock_guard<std::mutex> lk(g_servers_mu);
  if (add) g_servers.push_back(S);
  else g_servers.erase(std::remove(g_servers.begin(), g_servers.end(), S), g_servers.end());
}

// marks a request failed and wakes waiters; caller holds S->mu
void fail_locked(l2g_server_s* S, ServerRequest* r, int code, const std::string& msg) {
  r->state = L2G_REQ_FAILED;
  r->error = code;
  r->message = msg;
  --S->inflight;
  S->cv.notify_all();
}

// returns the arena ranges and slot of a request; caller holds S->mu
void release_locked(l2g_server_s* S, ServerRequest* r) {
  S->fl_pair.release(r->pair_off, r->npair);
  S->fl_prim.release(r->prim_off, r->nprimpair);
  S->fl_mat.release(r->mat_off, r->nmat);
  r->pair_off = r->prim_off = r->mat_off = -1;
  if (r->slot >= 0) S->free_slots.push_back(r->slot);
  r->slot = -1;
}

// slot and arena ranges for a prepared request; caller holds S->mu
bool reserve_locked(l2g_server_s* S, ServerRequest* r) {
  if (S->free_slots.empty()) return false;
  const int64_t a = S->fl_pair.alloc(r->npair);
  if (a < 0) return false;
  const int64_t b = S->fl_prim.alloc(r->nprimpair);
  if (b < 0) {
    S->fl_pair.release(a, r->npair);
    return false;
  }
  const int64_t c = S->fl_mat.alloc(r->nmat);
  if (c < 0) {
    S->fl_pair.release(a, r->npair);
    S->fl_prim.release(b, r->nprimpair);
    return false;
  }
  r->pair_off = a, r->prim_off = b, r->mat_off = c;
  r->slot = S->free_slots.back();
  S->free_slots.pop_back();
  return true;
}

// host and Schwarz set-up of one request (admission thread, no lock held)
void prepare(l2g_server_s* S, ServerRequest* r) {
  const int n = (int)r->l.size();
  int rc = l2g_basis_create(S->admit, n, r->l.data(), r->pure.data(), r->nprim.data(), r->alpha.data(),
                            r->coeff.data(), r->origin.data(), &r->basis);
  if (rc) throw L2gError(rc, g_err);
  fock_setup(r->setup, r->basis, S->cfg.pair_threshold, S->cfg.max_engine_precision, S->cfg.screening);
  fock_enumerate(r->setup, r->D.data(), r->precision, r->fp32, r->work);
  r->npair = (int64_t)r->setup.first.size();
  r->nprimpair = (int64_t)r->setup.P.size();
  r->nmat = (int64_t)r->nbf * r->nbf;
  r->left = 0;  // work items: deep quartets may be split into several
  for (const auto& it : r->work.cls) r->left += (int64_t)it.size();
  r->cursor.assign(r->work.cls.size(), 0);
}

// copies a reserved request's pair data and density into the arenas
// (admission stream) and rebases its work items onto the arena
void upload_request(l2g_server_s* S, ServerRequest* r) {
  hipStream_t st = S->admit->stream;
  const FockSetup& f = r->setup;
  std::vector<int> first(f.first);
  for (int& x : first) x += (int)r->prim_off;
  auto put = [&](DevBuf& arena, const void* src, size_t bytes, int64_t elem_off, size_t elem) {
    if (bytes) HIPCHK(hipMemcpyAsync(arena.as<char>() + elem_off * elem, src, bytes, hipMemcpyHostToDevice, st));
  };
  put(S->a_first, first.data(), first.size() * sizeof(int), r->pair_off, sizeof(int));
  put(S->a_np, f.np.data(), f.np.size() * sizeof(int), r->pair_off, sizeof(int));
  put(S->a_AB, f.AB.data(), f.AB.size() * sizeof(double), r->pair_off, 3 * sizeof(double));
  put(S->a_P, f.P.data(), f.P.size() * sizeof(double4), r->prim_off, sizeof(double4));
  put(S->a_PA, f.PA.data(), f.PA.size() * sizeof(double4), r->prim_off, sizeof(double4));
  put(S->a_aux, f.aux.data(), f.aux.size() * sizeof(double2), r->prim_off, sizeof(double2));
  put(S->a_ns, f.ns.data(), f.ns.size() * sizeof(double), r->prim_off, sizeof(double));
  put(S->a_dens, r->D.data(), r->D.size() * sizeof(double), r->mat_off, sizeof(double));
  if (r->nmat) {
    HIPCHK(hipMemsetAsync(S->a_jg.as<double>() + r->mat_off, 0, r->nmat * sizeof(double), st));
    HIPCHK(hipMemsetAsync(S->a_kg.as<double>() + r->mat_off, 0, r->nmat * sizeof(double), st));
    if (S->fx_hoff) {  // the high-word plane of the deterministic digest
      HIPCHK(hipMemsetAsync(S->a_jg.as<double>() + S->fx_hoff + r->mat_off, 0, r->nmat * sizeof(double), st));
      HIPCHK(hipMemsetAsync(S->a_kg.as<double>() + S->fx_hoff + r->mat_off, 0, r->nmat * sizeof(double), st));
    }
  }
  for (auto& it : r->work.cls) {
    for (int& x : it.bra) x += (int)r->pair_off;
    for (int& x : it.ket) x += (int)r->pair_off;
  }
  HIPCHK(hipHostMalloc((void**)&r->result, std::max<size_t>(3 * (size_t)r->nmat, 1) * sizeof(double)));
  HIPCHK(hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming));
  HIPCHK(hipStreamSynchronize(st));
}

// Admission worker: prepares the oldest unclaimed request (several workers
// prepare in parallel), then admits requests strictly in submission order,
// each as soon as a slot and arena space are free.
void admit_loop(l2g_server_s* S) {
  (void)hipSetDevice(S->cfg.device);
  for (;;) {
    ServerRequest* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto unclaimed = [&] {
        for (ServerRequest* q : S->pending)
          if (!q->claimed) return q;
        return (ServerRequest*)nullptr;
      };
      S->cv.wait(lk, [&] { return (S->stop && S->pending.empty()) || unclaimed(); });
      r = unclaimed();
      if (!r) return;
      r->claimed = true;
    }
    int rc = guarded([&] { prepare(S, r); });
    bool ok = rc == 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] { return S->pending.front() == r; });  // admission in submission order
      if (!ok) {
        fail_locked(S, r, rc, g_err);
      } else {
        // wait for a slot and arena space; a request that cannot fit an
        // empty server fails instead of waiting forever
        while (!reserve_locked(S, r)) {
          if (S->active.empty() && S->finishing.empty() && S->uploading == 0) {
            ok = false;
            fail_locked(S, r, L2G_E_OOM, "request does not fit the server's device memory or slots");
            break;
          }
          S->cv.wait(lk);
        }
      }
      S->pending.pop_front();
      if (ok) ++S->uploading;
      S->cv.notify_all();
    }
    if (!ok) {
      if (r->basis) l2g_basis_destroy(r->basis), r->basis = nullptr;
      continue;
    }
    rc = guarded([&] { upload_request(S, r); });
    std::unique_lock<std::mutex> lk(S->mu);
    --S->uploading;
    if (rc) {
      release_locked(S, r);
      fail_locked(S, r, rc, g_err);
      continue;
    }
    r->state = L2G_REQ_ACTIVE;
    r->t_admit = Clock::now();
    S->active.push_back(r);
    S->cv.notify_all();
  }
}

// moves up to k of r's remaining items, class by class, into the round
int64_t take(ServerRequest* r, int64_t k, std::vector<FockItems>& round) {
  int64_t got = 0;
  for (size_t c = 0; c < r->work.cls.size() && got < k; ++c) {
    const FockItems& it = r->work.cls[c];
    size_t& cur = r->cursor[c];
    if (cur >= it.size()) continue;
    const size_t e = std::min(it.size(), cur + (size_t)(k - got));
    FockItems& dst = round[c];
    dst.append(it, cur, e);
    dst.req.insert(dst.req.end(), e - cur, r->slot);
    got += (int64_t)(e - cur);
    cur = e;
  }
  r->left -= got;
  return got;
}

// retires requests whose finish step completed; caller holds S->mu
void retire_locked(l2g_server_s* S) {
  for (size_t i = 0; i < S->finishing.size();) {
    ServerRequest* r = S->finishing[i];
    const hipError_t q = hipEventQuery(r->done_ev);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    S->finishing.erase(S->finishing.begin() + i);
    release_locked(S, r);
    if (r->basis) l2g_basis_destroy(r->basis), r->basis = nullptr;
    r->setup = FockSetup();
    r->work = FockWork();
    r->D.clear();
    r->D.shrink_to_fit();
    if (q != hipSuccess) {
      fail_locked(S, r, L2G_E_LAUNCH, std::string("Fock kernels failed: ") + hipGetErrorString(q));
      continue;
    }
    r->state = L2G_REQ_DONE;
    r->t_done = Clock::now();
    --S->inflight;
    S->cv.notify_all();
  }
}

// one launch round over the admitted requests (launch thread)
void run_round(l2g_server_s* S, const std::vector<ServerRequest*>& act) {
  l2g_context_s* ctx = S->exec;
  hipStream_t st = ctx->stream;
  const int b = S->rnext;
  S->rnext ^= 1;
  HIPCHK(hipEventSynchronize(S->rev[b]));  // the round that last used buffer b is done
  std::vector<FockItems> round(nfock_classes());
  const int64_t budget = S->cfg.round_quartets;
  const int64_t share = std::max<int64_t>(1, budget / (int64_t)act.size());
  int64_t taken = 0;
  int nreq = 0;
  for (ServerRequest* r : act) {
    const int64_t t = take(r, std::min(share, r->left), round);
    taken += t;
    nreq += t > 0;
  }
  for (ServerRequest* r : act)  // leftover capacity, oldest first
    if (taken < budget && r->left > 0) taken += take(r, budget - taken, round);
  std::vector<L2gMatDesc>& mh = S->rmats_host[b];
  mh.assign(S->cfg.max_active, L2gMatDesc{0, 0});
  for (ServerRequest* r : act) mh[r->slot] = L2gMatDesc{r->mat_off, r->nbf};
  upload(S->rmats[b], mh, st);
  S->rbuf[b].upload_from(round, st);
  L2gEriArgs A = fock_args(ctx, S->cfg.screening, S->a_first.as<int>(), S->a_np.as<int>(), S->a_AB.as<double>(),
                           S->a_P.as<double4>(), S->a_PA.as<double4>(), S->a_aux.as<double2>(), S->a_ns.as<double>(),
                           S->a_dens.as<double>(), S->a_jg.as<double>(), S->a_kg.as<double>(),
                           S->rmats[b].as<L2gMatDesc>(), S->ybuf.as<double>(), S->ystride, S->fx_hoff);
  const ItemsView rv = S->rbuf[b].view();
  std::vector<int64_t> lanes(rv.count.size(), 0), yd(rv.count.size(), 0);
  for (size_t c = 0; c < rv.count.size(); ++c) {
    if (rv.count[c] == 0) continue;
    lanes[c] = fock_is_generic((int)c) ? rv.count[c] * 64 : rv.count[c];
    yd[c] = fock_ybuf((int)c);
  }
  std::vector<ClassSlot> slots;  // within the server's scratch (sized for the full grid)
  (void)plan_slots(slots, lanes, yd, ctx->ncu, S->block, l2g_context_s::kSide,
                   (int64_t)(S->ybuf.bytes / sizeof(double)));
  fock_launch_all(ctx, A, rv, true, S->grid, S->block, st, nullptr, &slots);
  HIPCHK(hipEventRecord(S->rev[b], st));
  // requests whose last quartets are in this round: finish on the device and
  // copy J, K, G to their pinned host buffers
  std::vector<ServerRequest*> done;
  for (ServerRequest* r : act) {
    if (r->left > 0) continue;
    const size_t NN = (size_t)r->nmat;
    double* jg = S->a_jg.as<double>() + r->mat_off;
    double* kg = S->a_kg.as<double>() + r->mat_off;
    double* g = S->a_dens.as<double>() + r->mat_off;
    launch_fock_finish(jg, kg, g, r->nbf, S->fx_hoff, st);
    if (NN) {
      HIPCHK(hipMemcpyAsync(r->result, jg, NN * sizeof(double), hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(r->result + NN, kg, NN * sizeof(double), hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(r->result + 2 * NN, g, NN * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipEventRecord(r->done_ev, st));
    done.push_back(r);
  }
  std::lock_guard<std::mutex> lk(S->mu);
  ++S->rounds;
  S->max_per_round = std::max<int64_t>(S->max_per_round, nreq);
  S->nq += taken;
  for (ServerRequest* r : done) {
    S->active.erase(std::find(S->active.begin(), S->active.end(), r));
    S->finishing.push_back(r);
  }
}

void exec_loop(l2g_server_s* S) {
  (void)hipSetDevice(S->cfg.device);
  for (;;) {
    std::vector<ServerRequest*> act;
    ServerRequest* wait_for = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      retire_locked(S);
      if (S->active.empty()) {
        if (!S->finishing.empty()) {
          wait_for = S->finishing.front();
        } else {
          if (S->stop && S->inflight == 0) return;
          S->cv.wait(lk, [&] { return !S->active.empty() || (S->stop && S->inflight == 0); });
          continue;
        }
      } else {
        act = S->active;
      }
    }
    if (wait_for) {
      (void)hipEventSynchronize(wait_for->done_ev);
      continue;
    }
    const int rc = guarded([&] { run_round(S, act); });
    if (rc) {
      // a launch failed: every request of the round fails with it.  Kernels
      // of the round queued before the failure may still be adding into the
      // requests' accumulators: wait for the exec streams (the side streams
      // were joined into the main one) before their arena ranges go back to
      // the free lists, where the next admission zeroes and reuses them on
      // another stream
      (void)hipStreamSynchronize(S->exec->stream);
      for (int k = 0; k < l2g_context_s::kSide; ++k) (void)hipStreamSynchronize(S->exec->side[k]);
      std::lock_guard<std::mutex> lk(S->mu);
      for (ServerRequest* r : act) {
        auto it = std::find(S->active.begin(), S->active.end(), r);
        if (it == S->active.end()) continue;
        S->active.erase(it);
        release_locked(S, r);
        fail_locked(S, r, rc, g_err);
      }
    }
  }
}

}  // namespace

extern "C" {

int l2g_server_create(const l2g_server_config* config, l2g_server* out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("l2g_server_create: null argument");
    auto S = std::make_unique<l2g_server_s>();
    l2g_server_config c{};
    if (config) c = *config;
    if (c.hbm_bytes <= 0) c.hbm_bytes = 16.0 * (1ull << 30);
    if (c.max_active <= 0) c.max_active = 16;
    if (c.round_quartets <= 0) c.round_quartets = 1 << 20;
    if (c.screening == 0) c.screening = L2G_SCREEN_SCHWARZINF;
    if (c.pair_threshold <= 0) c.pair_threshold = 1e-12;
    if (c.max_engine_precision <= 0) c.max_engine_precision = std::numeric_limits<double>::epsilon() / 1e10;
    S->cfg = c;
    int rc = l2g_init(c.device, &S->exec);
    if (rc) throw L2gError(rc, g_err);
    rc = l2g_init(c.device, &S->admit);
    if (rc) {
      l2g_finalize(S->exec);
      throw L2gError(rc, g_err);
    }
    // arenas: 60% matrices (D, jg, kg: 24 bytes per element; 40 with the
    // deterministic digest's second plane of jg and kg), 35% primitive
    // pairs (88 bytes), 5% shell pairs (32 bytes)
    const double hb = c.hbm_bytes;
    const int planes = c.deterministic ? 2 : 1;
    const int64_t cap_mat = (int64_t)(0.60 * hb / (8 + 16 * planes)), cap_prim = (int64_t)(0.35 * hb / 88),
                  cap_pair = (int64_t)(0.05 * hb / 32);
    S->fl_mat.init(cap_mat);
    S->fl_prim.init(cap_prim);
    S->fl_pair.init(cap_pair);
    S->fx_hoff = c.deterministic ? cap_mat : 0;
    S->a_dens.alloc(cap_mat * sizeof(double));
    S->a_jg.alloc(planes * cap_mat * sizeof(double));
    S->a_kg.alloc(planes * cap_mat * sizeof(double));
    S->a_P.alloc(cap_prim * sizeof(double4));
    S->a_PA.alloc(cap_prim * sizeof(double4));
    S->a_aux.alloc(cap_prim * sizeof(double2));
    S->a_ns.alloc(cap_prim * sizeof(double));
    S->a_first.alloc(cap_pair * sizeof(int));
    S->a_np.alloc(cap_pair * sizeof(int));
    S->a_AB.alloc(cap_pair * 3 * sizeof(double));
    int64_t ymax = 1;
    for (int k = 0; k < nfock_classes(); ++k) ymax = std::max<int64_t>(ymax, fock_ybuf(k));
    S->grid = S->exec->ncu * 8;
    S->ystride = (int64_t)S->grid * S->block;
    S->ybuf.alloc(ymax * S->ystride * sizeof(double));
    for (int k = 0; k < 2; ++k) HIPCHK(hipEventCreateWithFlags(&S->rev[k], hipEventDisableTiming));
    for (int k = c.max_active - 1; k >= 0; --k) S->free_slots.push_back(k);
    l2g_server_s* p = S.get();
    for (int k = 0; k < 4; ++k) p->admit_thr.emplace_back(admit_loop, p);
    p->exec_thr = std::thread(exec_loop, p);
    server_registry(p, true);
    *out = S.release();
  });
}

int l2g_request_submit(l2g_server S, int nshell, const int* l, const int* pure, const int* nprim, const double* alpha,
                       const double* coeff, const double* origin, const double* D, double precision,
                       double fp32_threshold, int64_t* id) {
  return guarded([&] {
    if (!S || !id || nshell < 0 || (nshell > 0 && (!l || !pure || !nprim || !alpha || !coeff || !origin || !D)))
      throw std::invalid_argument("l2g_request_submit: bad arguments");
    if (!(fp32_threshold >= 0)) throw std::invalid_argument("fp32_threshold must be >= 0");
    auto r = std::make_unique<ServerRequest>();
    r->l.assign(l, l + nshell);
    r->pure.assign(pure, pure + nshell);
    r->nprim.assign(nprim, nprim + nshell);
    int64_t np = 0, nbf = 0;
    for (int s = 0; s < nshell; ++s) {
      if (nprim[s] <= 0 || l[s] < 0) throw std::invalid_argument("l2g_request_submit: bad shell");
      np += nprim[s];
      nbf += nfun(l[s], pure[s]);
    }
    r->alpha.assign(alpha, alpha + np);
    r->coeff.assign(coeff, coeff + np);
    r->origin.assign(origin, origin + 3 * (size_t)nshell);
    r->D.assign(D, D + nbf * nbf);
    r->nbf = (int)nbf;
    r->precision = precision;
    r->fp32 = fp32_threshold;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) throw L2gError(L2G_E_STATE, "server is shutting down");
    r->id = S->next_id++;
    *id = r->id;
    S->pending.push_back(r.get());
    S->reqs.emplace(r->id, std::move(r));
    ++S->inflight;
    S->cv.notify_all();
  });
}

int l2g_request_state(l2g_server S, int64_t id, int* state) {
  return guarded([&] {
    if (!S || !state) throw std::invalid_argument("null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) throw std::invalid_argument("unknown request id");
    *state = it->second->state;
  });
}

int l2g_request_wait(l2g_server S, int64_t id) {
  int code = 0;
  std::string msg;
  const int rc = guarded([&] {
    if (!S) throw std::invalid_argument("null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) throw std::invalid_argument("unknown request id");
    ServerRequest* r = it->second.get();
    S->cv.wait(lk, [&] { return r->state == L2G_REQ_DONE || r->state == L2G_REQ_FAILED; });
    if (r->state == L2G_REQ_FAILED) {
      code = r->error;
      msg = r->message;
    }
  });
  if (rc) return rc;
  if (code) {
    g_err = msg;
    return code;
  }
  return L2G_OK;
}

int l2g_request_result(l2g_server S, int64_t id, double* J, double* K, double* G, double* ms) {
  return guarded([&] {
    if (!S) throw std::invalid_argument("null server");
    std::unique_ptr<ServerRequest> r;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      auto it = S->reqs.find(id);
      if (it == S->reqs.end()) throw std::invalid_argument("unknown request id");
      if (it->second->state != L2G_REQ_DONE) throw L2gError(L2G_E_STATE, "request is not done");
      r = std::move(it->second);
      S->reqs.erase(it);
    }
    const size_t NN = (size_t)r->nmat;
    if (J) std::memcpy(J, r->result, NN * sizeof(double));
    if (K) std::memcpy(K, r->result + NN, NN * sizeof(double));
    if (G) std::memcpy(G, r->result + 2 * NN, NN * sizeof(double));
    if (ms) {
      ms[0] = std::chrono::duration<double, std::milli>(r->t_admit - r->t_submit).count();
      ms[1] = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    }
    (void)hipHostFree(r->result);
    r->result = nullptr;
    if (r->done_ev) (void)hipEventDestroy(r->done_ev);
  });
}

int l2g_server_stats(l2g_server S, int64_t* rounds, int64_t* max_requests_per_round, int64_t* nquartets) {
  return guarded([&] {
    if (!S) throw std::invalid_argument("null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (rounds) *rounds = S->rounds;
    if (max_requests_per_round) *max_requests_per_round = S->max_per_round;
    if (nquartets) *nquartets = S->nq;
  });
}

// At process exit: the threads of every server not destroyed by its owner
// finish the requests in flight and stop, before the HIP runtime goes.
extern "C++" void l2g_impl::stop_servers_at_exit() {
  std::vector<l2g_server_s*> live;
  {
    std::lock_guard<std::mutex> lk(g_servers_mu);
    live.swap(g_servers);
  }
  for (l2g_server_s* S : live) {
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
      S->cv.notify_all();
    }
    for (auto& t : S->admit_thr)
      if (t.joinable()) t.join();
    if (S->exec_thr.joinable()) S->exec_thr.join();
  }
}

int l2g_server_destroy(l2g_server S) {
  if (!S || g_exiting) return L2G_OK;
  server_registry(S, false);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv.notify_all();
  }
  for (auto& t : S->admit_thr)
    if (t.joinable()) t.join();
  if (S->exec_thr.joinable()) S->exec_thr.join();
  (void)hipSetDevice(S->cfg.device);
  (void)hipStreamSynchronize(S->exec->stream);
  for (auto& kv : S->reqs) {
    ServerRequest* r = kv.second.get();
    if (r->result) (void)hipHostFree(r->result);
    if (r->done_ev) (void)hipEventDestroy(r->done_ev);
    if (r->basis) l2g_basis_destroy(r->basis);
  }
  for (auto& e : S->rev)
    if (e) (void)hipEventDestroy(e);
  l2g_context ex = S->exec, ad = S->admit;
  delete S;  // frees the arenas before the contexts go
  l2g_finalize(ex);
  l2g_finalize(ad);
  return L2G_OK;
}

}  // extern "C"
