// Resident single-quartet evaluator: the per-call path behind the drop-in
// Engine::compute(s1, s2, s3, s4) / compute2<coulomb, xx_xx, 0>(..., spbra,
// spket) (include/libint2/engine.h:747-791, engine.impl.h:1151-2113).
//
// libint's callers loop over shell quartets and call the Engine once per
// quartet, so what matters on this path is the fixed cost of one call.  A
// quartet engine holds, from its creation on, every device buffer one
// quartet needs (pair data for max_nprim^2 primitive pairs per side, the
// output, the lane scratch of the largest class) and pinned host staging.
// A call builds the two shell pairs on the host, packs the work item and
// pair data into the staging buffer, and makes one host-to-device copy,
// one kernel launch per operator pass (one block; the class's register
// kernel, or the generic kernel for a class without one) and one
// device-to-host copy of the integrals and the screened-out flag.
#include <limits>

#include "l2g_internal.h"

using namespace l2g_impl;

struct l2g_quartet_engine_s {
  l2g_context_s* ctx = nullptr;
  DevBuf dbuf, ybuf;  // staged inputs, then the flag slot and the integrals; lane scratch
  QuartetBuffers qb;
  ~l2g_quartet_engine_s() {
    if (g_exiting) return;
    if (qb.hin) (void)hipHostFree(qb.hin);
    if (qb.hout) (void)hipHostFree(qb.hout);
    if (ctx) --ctx->live;
  }
};

namespace l2g_impl {

int64_t quartet_ylane(int max_l) {
  int64_t y = 0;
  for (const L2gClassInfo* tab : {l2g_gen_classes, l2g_gen_classes_erf}) {
    const int n = tab == l2g_gen_classes ? l2g_gen_nclasses : l2g_gen_nclasses_erf;
    for (int i = 0; i < n; ++i) {
      const L2gClassInfo& c = tab[i];
      if (std::max(std::max(c.la, c.lb), std::max(c.lc, c.ld)) <= max_l) y = std::max<int64_t>(y, c.ybuf_doubles);
    }
  }
  return y;
}

}  // namespace l2g_impl

extern "C" {

int l2g_quartet_engine_create(l2g_context ctx, int max_nprim, int max_l, l2g_quartet_engine* out) {
  return guarded([&] {
    if (!ctx) throw L2gError(L2G_E_STATE, "null context");
    if (!out || max_nprim <= 0 || max_l < 0) throw std::invalid_argument("l2g_quartet_engine_create: bad arguments");
    if (max_l > lmax_supported())
      throw L2gError(L2G_E_LMAX, "max_l " + std::to_string(max_l) + " exceeds the compiled limit " +
                                     std::to_string(lmax_supported()));
    HIPCHK(hipSetDevice(ctx->device));
    auto e = std::make_unique<l2g_quartet_engine_s>();
    QuartetBuffers& qb = e->qb;
    qb.max_nprim = max_nprim;
    qb.max_l = max_l;
    qb.in_bytes = QuartetLayout(2 * (size_t)max_nprim * (size_t)max_nprim).total;
    const size_t nc = (size_t)ncart(max_l);
    qb.out_doubles = nc * nc * nc * nc;
    qb.ylane = quartet_ylane(max_l);
    // device block: inputs (at most in_bytes), one double for the
    // screened-out flag, the integrals
    e->dbuf.alloc(qb.in_bytes + (qb.out_doubles + 1) * sizeof(double));
    e->ybuf.alloc(std::max<int64_t>(qb.ylane, 1) * kQuartetBlock * sizeof(double));
    qb.d = e->dbuf.as<unsigned char>();
    qb.ybuf = e->ybuf.as<double>();
    HIPCHK(hipHostMalloc((void**)&qb.hin, qb.in_bytes + sizeof(double), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&qb.hout, (qb.out_doubles + 1) * sizeof(double), hipHostMallocDefault));
    e->ctx = ctx;
    ++ctx->live;
    *out = e.release();
  });
}

int l2g_quartet_engine_destroy(l2g_quartet_engine e) {
  if (e && !g_exiting) {
    (void)hipSetDevice(e->ctx->device);
    (void)hipStreamSynchronize(e->ctx->stream);
  }
  if (!g_exiting) delete e;
  return L2G_OK;
}

int l2g_quartet_compute(l2g_quartet_engine e, const int* l, const int* pure, const int* nprim, const double* alpha,
                        const double* coeff, const double* origin, const double* ln_scr, const double* nonsph,
                        double precision, int screening, double scale, int uniform_cart, int oper, double omega,
                        double* out, int8_t* nonzero) {
  return guarded([&] {
    if (!e) throw L2gError(L2G_E_STATE, "null quartet engine");
    quartet_compute_impl(e->ctx, e->qb, l, pure, nprim, alpha, coeff, origin, ln_scr, nonsph, precision, screening,
                         scale, uniform_cart, oper, omega, out, nonzero);
  });
}

}  // extern "C"

namespace l2g_impl {

void quartet_compute_impl(l2g_context_s* ctx, const QuartetBuffers& qb, const int* l, const int* pure,
                          const int* nprim, const double* alpha, const double* coeff, const double* origin,
                          const double* ln_scr, const double* nonsph, double precision, int screening, double scale,
                          int uniform_cart, int oper, double omega, double* out, int8_t* nonzero) {
  if (!l || !pure || !nprim || !alpha || !coeff || !origin || !out)
    throw std::invalid_argument("l2g_quartet_compute: null argument");
  if (oper < 0 || oper > 2 || !(omega >= 0)) throw std::invalid_argument("l2g_quartet_compute: bad operator or omega");
  const bool given = ln_scr != nullptr;
  if (!given && screening != L2G_SCREEN_ORIGINAL && screening != L2G_SCREEN_CONSERVATIVE)
    throw std::invalid_argument(
        "without precomputed shell pair data only the Original or Conservative screening methods are available");
  if (screening != L2G_SCREEN_ORIGINAL && screening != L2G_SCREEN_CONSERVATIVE && screening != L2G_SCREEN_SCHWARZ &&
      screening != L2G_SCREEN_SCHWARZINF)
    throw std::invalid_argument("l2g_quartet_compute: unknown screening method");
  // the four shells as a host basis (no device data)
  l2g_basis_s B;
  B.ctx = nullptr;
  B.nshell = 4;
  B.l.assign(l, l + 4);
  B.pure.assign(pure, pure + 4);
  B.nprim.assign(nprim, nprim + 4);
  B.off.assign(5, 0);
  for (int s = 0; s < 4; ++s) {
    if (l[s] < 0) throw std::invalid_argument("negative angular momentum");
    if (l[s] > qb.max_l)
      throw L2gError(L2G_E_LMAX, "angular momentum " + std::to_string(l[s]) + " exceeds the " + qb.owner +
                                     "'s max_l " + std::to_string(qb.max_l));
    if (nprim[s] <= 0) throw std::invalid_argument("shell without primitives");
    if (nprim[s] > qb.max_nprim)
      throw std::invalid_argument("shell of " + std::to_string(nprim[s]) + " primitives exceeds the " + qb.owner +
                                  "'s max_nprim " + std::to_string(qb.max_nprim));
    B.off[s + 1] = B.off[s] + nprim[s];
  }
  const int np = B.off[4];
  B.alpha.assign(alpha, alpha + np);
  B.coeff.assign(coeff, coeff + np);
  B.lnc.resize(np);
  for (int p = 0; p < np; ++p) B.lnc[p] = std::log(std::fabs(B.coeff[p]));
  B.origin.assign(origin, origin + 12);
  double prec = 0, ln_prec = std::numeric_limits<double>::lowest();
  if (precision > 0) {
    prec = precision;
    ln_prec = std::log(precision);
  }
  // canonical order (engine.impl.h:1181-1208), as l2g_eri_plan_create
  const bool stb = B.swap_pair(0, 1), stk = B.swap_pair(2, 3), sbk = (l[0] + l[1]) > (l[2] + l[3]);
  const int b1 = sbk ? (stk ? 3 : 2) : (stb ? 1 : 0);
  const int b2 = sbk ? (stk ? 2 : 3) : (stb ? 0 : 1);
  const int k1 = sbk ? (stb ? 1 : 0) : (stk ? 3 : 2);
  const int k2 = sbk ? (stb ? 0 : 1) : (stk ? 2 : 3);
  PairBuilder pb(B, ln_prec, screening);
  if (given) {
    // bra pair (0, 1): n0 n1 values; ket pair (2, 3): n2 n3 values after it
    const size_t nb = (size_t)nprim[0] * nprim[1];
    pb.given.emplace((uint64_t)0 << 32 | 1u, PairBuilder::Given{ln_scr, nonsph});
    pb.given.emplace((uint64_t)2 << 32 | 3u, PairBuilder::Given{ln_scr + nb, nonsph ? nonsph + nb : nullptr});
  }
  const int bp = pb.get(b1, b2), kp = pb.get(k1, k2);
  const size_t npp = pb.P.size();
  auto purebit = [&](int sh) { return (pure[sh] && l[sh] > 0) ? 1 : 0; };
  const int meta = (stb ? 1 : 0) | (stk ? 2 : 0) | (sbk ? 4 : 0) | (purebit(b1) << 3) | (purebit(b2) << 4) |
                   (purebit(k1) << 5) | (purebit(k2) << 6);
  const int64_t n = (int64_t)nfun(l[0], pure[0]) * nfun(l[1], pure[1]) * nfun(l[2], pure[2]) * nfun(l[3], pure[3]);
  const int cls = ctx->force_generic ? -1 : ctx->class_of[l[b1]][l[b2]][l[k1]][l[k2]];
  const int gcls = ctx->generic_of[l[b1]][l[b2]][l[k1]][l[k2]];
  if (cls < 0 && gcls < 0) throw L2gError(L2G_E_LMAX, "no kernel for this angular momentum class");
  // stage the inputs
  const QuartetLayout L(npp);
  if (L.total > qb.in_bytes || (size_t)n > qb.out_doubles)
    throw std::logic_error(std::string(qb.owner) + ": staging buffer too small");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  HIPCHK(hipStreamSynchronize(st));  // the previous call's copies are done with the staging buffers
  unsigned char* h = qb.hin;
  int* hi = reinterpret_cast<int*>(h + L.ints);
  hi[0] = bp, hi[1] = kp, hi[2] = meta;
  hi[3] = pb.first[0], hi[4] = pb.first[1], hi[5] = pb.np[0], hi[6] = pb.np[1], hi[7] = 0;
  long long* hl = reinterpret_cast<long long*>(h + L.lls);
  hl[0] = 1, hl[1] = 0;  // output after the flag slot; caller index 0
  std::memcpy(h + L.ab, pb.AB.data(), 6 * sizeof(double));
  std::memcpy(h + L.P, pb.P.data(), npp * sizeof(double4));
  std::memcpy(h + L.PA, pb.PA.data(), npp * sizeof(double4));
  std::memcpy(h + L.aux, pb.aux.data(), npp * sizeof(double2));
  std::memcpy(h + L.ns, pb.ns.data(), npp * sizeof(double));
  std::memset(h + L.total, 0, sizeof(double));  // the flag slot starts at 0 (the kernel writes it)
  unsigned char* d = qb.d;
  HIPCHK(hipMemcpyAsync(d, h, L.total + sizeof(double), hipMemcpyHostToDevice, st));
  double* dout = reinterpret_cast<double*>(d + L.total);  // [0] flag, [1..n] integrals
  L2gEriArgs A;
  std::memset(&A, 0, sizeof(A));
  const int* di = reinterpret_cast<const int*>(d + L.ints);
  A.bra_pair = di;
  A.ket_pair = di + 1;
  A.meta = di + 2;
  A.pair_first = di + 3;
  A.pair_np = di + 5;
  A.out_off = reinterpret_cast<const long long*>(d + L.lls);
  A.qidx = reinterpret_cast<const long long*>(d + L.lls) + 1;
  A.pair_AB = reinterpret_cast<const double*>(d + L.ab);
  A.pp_P = reinterpret_cast<const double4*>(d + L.P);
  A.pp_PA = reinterpret_cast<const double4*>(d + L.PA);
  A.pp_aux = reinterpret_cast<const double2*>(d + L.aux);
  A.pp_nonsph = reinterpret_cast<const double*>(d + L.ns);
  A.n = 1;
  A.boys = ctx->boys.as<double>();
  A.precision = prec;
  A.ln_precision = ln_prec;
  A.scale = scale;
  A.method = screening;
  A.shg_gaussian = ctx->shg_gaussian;
  A.uniform_cart = uniform_cart ? 1 : 0;
  A.out = dout;
  A.nz = reinterpret_cast<signed char*>(dout);  // the flag slot (qidx 0)
  A.ybuf = qb.ybuf;
  A.ystride = kQuartetBlock;
  // operator passes (attenuation, subtract), as l2g_eri_plan_execute
  std::vector<std::pair<double, int>> passes;
  if (oper == 0) passes = {{0.0, 0}};
  else if (oper == 1) passes = omega > 0 ? std::vector<std::pair<double, int>>{{omega, 0}}
                                         : std::vector<std::pair<double, int>>{{0.0, 0}, {0.0, 1}};
  else passes = omega > 0 ? std::vector<std::pair<double, int>>{{0.0, 0}, {omega, 1}}
                          : std::vector<std::pair<double, int>>{{0.0, 0}};
  for (const auto& ps : passes) {
    A.omega = ps.first;
    A.store_sub = ps.second;
    int erf_idx = -1;
    if (A.omega > 0 && cls >= 0) erf_idx = ctx->classerf_of[l[b1]][l[b2]][l[k1]][l[k2]];
    if (erf_idx >= 0) {
      l2g_gen_classes_erf[erf_idx].launch(A, 1, kQuartetBlock, st);
      LAUNCHCHK("erf ERI kernel (quartet engine)");
    } else if (cls < 0 || A.omega > 0) {
      if (gcls < 0) throw L2gError(L2G_E_LMAX, "no generic-kernel program for this class");
      launch_generic(kGenericEri, A, l2g_gen_generic[gcls], ctx->gprog.as<int4>(), ctx->gidx.as<int>(), 1, st);
    } else {
      l2g_gen_classes[cls].launch(A, 1, kQuartetBlock, st);
      LAUNCHCHK("ERI kernel (quartet engine)");
    }
  }
  HIPCHK(hipMemcpyAsync(qb.hout, dout, (size_t)(n + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::memcpy(out, qb.hout + 1, (size_t)n * sizeof(double));
  if (nonzero) *nonzero = *reinterpret_cast<const signed char*>(qb.hout) ? 1 : 0;
}

}  // namespace l2g_impl
