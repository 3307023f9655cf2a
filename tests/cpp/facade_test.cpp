// Drives the libint2-compatible façade (include/libint2_amd/engine.h) the way
// libint's own tests drive Engine (tests/unit/test-core.cc:59-77,
// test-2body.cc): prints every integral of a few quartets for comparison
// with the oracle, and checks the error behaviour.
//   facade_test cpu   -> only the checks that need no GPU
//   facade_test       -> everything (needs a GPU)
#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <libint2_amd/engine.h>

using libint2::BraKet;
using libint2::Engine;
using libint2::Operator;
using libint2::Shell;

static int fails = 0;
#define CHECK(x)                                              \
  do {                                                        \
    if (!(x)) {                                               \
      std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #x, __LINE__); \
      ++fails;                                                \
    }                                                         \
  } while (0)

template <class E, class F>
static bool throws(F&& f) {
  try {
    f();
  } catch (const E&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

using libint2::ShellPair;
using libint2::ScreeningMethod;

// The basis of the Fock-loop check: three atoms, per atom s (3 primitives),
// p (2) and pure d (1).  Written out in the Python test as well.
static std::vector<Shell> fock_basis() {
  const double xyz[3][3] = {{0.0, 0.0, 0.0}, {1.4, 0.3, -0.2}, {-0.5, 1.6, 0.4}};
  std::vector<Shell> b;
  for (auto& o : xyz) {
    std::array<double, 3> O{{o[0], o[1], o[2]}};
    b.push_back(Shell{{5.0, 1.2, 0.3}, {{0, false, {0.3, 0.5, 0.4}}}, O});
    b.push_back(Shell{{2.1, 0.45}, {{1, false, {0.6, 0.5}}}, O});
    b.push_back(Shell{{0.8}, {{2, true, {1.0}}}, O});
  }
  return b;
}

// compute_shellpairs + compute_schwarz_ints + compute_2body_fock of
// tests/hartree-fock/hartree-fock++.cc:1231-1436, 1574-1772, single-threaded,
// through the façade: SchwarzInf primitive screening with ShellPairs built by
// a Schwarz evaluator over primitive shells, and compute2 with those pairs.
// Returns G for the density D (row-major nbf x nbf).
static std::vector<double> fock_loop(const std::vector<Shell>& obs, const std::vector<double>& D, size_t nbf,
                                     double fock_precision) {
  const auto max_engine_precision = std::numeric_limits<double>::epsilon() / 1e10;
  const size_t nsh = obs.size();
  std::vector<size_t> s2bf(nsh);
  for (size_t i = 0, f = 0; i < nsh; ++i) s2bf[i] = f, f += obs[i].size();
  // significant pairs (s2 <= s1): same centre or overlap norm >= 1e-12
  Engine ov(Operator::overlap, 3, 2, 0);
  ov.set_precision(0.);
  Engine sch(Operator::coulomb, 3, 2, 0, 0.0);
  std::map<size_t, std::vector<size_t>> splist;
  std::map<size_t, std::vector<std::shared_ptr<ShellPair>>> spdata;
  auto schwarz_factor = [&](const Shell& s1, size_t p1, const Shell& s2, size_t p2) -> double {
    auto ps1 = s1.extract_primitive(p1, false);
    auto ps2 = s2.extract_primitive(p2, false);
    const size_t n12 = ps1.size() * ps2.size();
    sch.compute(ps1, ps2, ps1, ps2);
    if (!sch.results()[0]) return 0.;
    double m = 0;
    for (size_t i = 0; i < n12 * n12; ++i) m = std::max(m, std::abs(sch.results()[0][i]));
    return std::sqrt(m);
  };
  for (size_t s1 = 0; s1 < nsh; ++s1)
    for (size_t s2 = 0; s2 <= s1; ++s2) {
      bool sig = obs[s1].O == obs[s2].O;
      if (!sig) {
        ov.compute(obs[s1], obs[s2]);
        double nrm = 0;
        for (size_t i = 0; i < obs[s1].size() * obs[s2].size(); ++i) nrm += ov.results()[0][i] * ov.results()[0][i];
        sig = std::sqrt(nrm) >= 1e-12;
      }
      if (!sig) continue;
      splist[s1].push_back(s2);
      spdata[s1].push_back(std::make_shared<ShellPair>(obs[s1], obs[s2], std::log(max_engine_precision),
                                                       ScreeningMethod::SchwarzInf, schwarz_factor));
    }
  // Schwarz matrix sqrt(max |(ab|ab)|), no primitive screening
  std::vector<double> K(nsh * nsh, 0.0);
  for (size_t s1 = 0; s1 < nsh; ++s1)
    for (size_t s2 = 0; s2 <= s1; ++s2) {
      sch.compute2<Operator::coulomb, BraKet::xx_xx, 0>(obs[s1], obs[s2], obs[s1], obs[s2]);
      double m = 0;
      const size_t n12 = obs[s1].size() * obs[s2].size();
      for (size_t i = 0; i < n12 * n12; ++i) m = std::max(m, std::abs(sch.results()[0][i]));
      K[s1 * nsh + s2] = K[s2 * nsh + s1] = std::sqrt(m);
    }
  // infinity norms of the density's shell blocks
  std::vector<double> Dn(nsh * nsh, 0.0);
  double Dmax = 0;
  for (size_t a = 0; a < nsh; ++a)
    for (size_t b = 0; b < nsh; ++b) {
      double m = 0;
      for (size_t i = 0; i < obs[a].size(); ++i)
        for (size_t j = 0; j < obs[b].size(); ++j) m = std::max(m, std::abs(D[(s2bf[a] + i) * nbf + s2bf[b] + j]));
      Dn[a * nsh + b] = m;
      Dmax = std::max(Dmax, m);
    }
  const double needed = fock_precision / Dmax;
  Engine engine(Operator::coulomb, 3, 2, 0);
  engine.set(ScreeningMethod::SchwarzInf);
  engine.set_precision(needed);
  const auto& buf = engine.results();
  std::vector<double> g(nbf * nbf, 0.0);
  for (size_t s1 = 0; s1 < nsh; ++s1) {
    auto sp12_it = spdata[s1].begin();
    for (size_t s2 : splist[s1]) {
      const ShellPair* sp12 = (sp12_it++)->get();
      const double Dn12 = Dn[s1 * nsh + s2];
      for (size_t s3 = 0; s3 <= s1; ++s3) {
        const double Dn123 = std::max(Dn[s1 * nsh + s3], std::max(Dn[s2 * nsh + s3], Dn12));
        auto sp34_it = spdata[s3].begin();
        const size_t s4_max = (s1 == s3) ? s2 : s3;
        for (size_t s4 : splist[s3]) {
          if (s4 > s4_max) break;
          const ShellPair* sp34 = (sp34_it++)->get();
          const double Dn1234 =
              std::max(Dn[s1 * nsh + s4], std::max(Dn[s2 * nsh + s4], std::max(Dn[s3 * nsh + s4], Dn123)));
          if (Dn1234 * K[s1 * nsh + s2] * K[s3 * nsh + s4] < fock_precision) continue;
          const double deg = (s1 == s2 ? 1 : 2) * (s3 == s4 ? 1 : 2) * (s1 == s3 ? (s2 == s4 ? 1 : 2) : 2);
          engine.set_precision(Dn1234 != 0. ? fock_precision / Dn1234 : needed);
          engine.compute2<Operator::coulomb, BraKet::xx_xx, 0>(obs[s1], obs[s2], obs[s3], obs[s4], sp12, sp34);
          if (!buf[0]) continue;
          const size_t n1 = obs[s1].size(), n2 = obs[s2].size(), n3 = obs[s3].size(), n4 = obs[s4].size();
          for (size_t f1 = 0, f = 0; f1 < n1; ++f1)
            for (size_t f2 = 0; f2 < n2; ++f2)
              for (size_t f3 = 0; f3 < n3; ++f3)
                for (size_t f4 = 0; f4 < n4; ++f4, ++f) {
                  const size_t a = s2bf[s1] + f1, b = s2bf[s2] + f2, c = s2bf[s3] + f3, d = s2bf[s4] + f4;
                  const double v = buf[0][f] * deg;
                  g[a * nbf + b] += D[c * nbf + d] * v;
                  g[c * nbf + d] += D[a * nbf + b] * v;
                  g[a * nbf + c] -= 0.25 * D[b * nbf + d] * v;
                  g[b * nbf + d] -= 0.25 * D[a * nbf + c] * v;
                  g[a * nbf + d] -= 0.25 * D[b * nbf + c] * v;
                  g[b * nbf + c] -= 0.25 * D[a * nbf + d] * v;
                }
        }
      }
    }
  }
  std::vector<double> G(nbf * nbf);
  for (size_t i = 0; i < nbf; ++i)
    for (size_t j = 0; j < nbf; ++j) G[i * nbf + j] = 0.5 * (g[i * nbf + j] + g[j * nbf + i]);
  return G;
}

int main(int argc, char** argv) {
  const bool cpu_only = argc > 1 && std::strcmp(argv[1], "cpu") == 0;
  // Shell normalization known answer (test-core.cc:39-57)
  Shell d1{{1.0}, {{2, false, {1.0}}}, {{0, 0, 0}}};
  CHECK(std::abs(d1.contr[0].coeff[0] - 1.64592278064949) < 1e-13);
  // default-initialized Engine misuse (test-core.cc:64-77)
  CHECK(throws<Engine::using_default_initialized>([] { Engine{}.set(Operator::coulomb); }));
  CHECK(throws<Engine::using_default_initialized>([] { Engine{}.set(BraKet::xx_xx); }));
  CHECK(!throws<std::exception>([] { Engine{}.prescale_by(1.5); }));
  // derivative orders beyond the first, and derivatives of 3-centre or
  // attenuated integrals, are refused before any GPU work
  CHECK(throws<std::invalid_argument>([] { Engine(Operator::coulomb, 2, 2, 2); }));
  CHECK(throws<std::invalid_argument>([] { Engine(Operator::erf_coulomb, 2, 2, 1, 1e-15, 0.5); }));
  // one-body operators go with BraKet::x_x only, at deriv_order 0
  CHECK(throws<std::invalid_argument>([] { Engine(Operator::overlap, 2, 2, 1); }));
  CHECK(throws<std::invalid_argument>([] { Engine(Operator::coulomb, 2, 2, 0, 1e-15, 0, BraKet::x_x); }));
  CHECK(throws<std::invalid_argument>([] { Engine(Operator::kinetic, 2, 2, 0, 1e-15, 0, BraKet::xx_xx); }));
  if (cpu_only) {
    std::printf("cpu checks: %s\n", fails ? "FAIL" : "ok");
    return fails ? 1 : 0;
  }
  libint2::initialize();
  CHECK(throws<Engine::lmax_exceeded>([] { Engine(Operator::coulomb, 3, l2g_max_l() + 1); }));
  // s p d shells, contracted, distinct centres
  Shell s{{0.9, 3.1}, {{0, false, {0.4, 0.7}}}, {{0.1, 0.2, 0.3}}};
  Shell p{{1.3, 0.4}, {{1, false, {0.5, 0.6}}}, {{1.0, 0.0, -0.2}}};
  Shell d{{2.2}, {{2, true, {1.0}}}, {{-0.3, 0.8, 0.1}}};
  Shell f0{{0.7, 1.9}, {{2, false, {0.3, 0.9}}}, {{0.4, -0.5, 0.6}}};
  std::vector<Shell> basis = {s, p, d, f0};
  Engine e(Operator::coulomb, 2, 2);
  const auto& buf = e.results();
  const int quartets[][4] = {{0, 1, 2, 3}, {2, 2, 1, 0}, {3, 2, 1, 1}, {1, 0, 3, 3}, {2, 3, 2, 3}};
  for (auto& q : quartets) {
    e.compute(basis[q[0]], basis[q[1]], basis[q[2]], basis[q[3]]);
    CHECK(buf[0] != nullptr);
    const size_t n = basis[q[0]].size() * basis[q[1]].size() * basis[q[2]].size() * basis[q[3]].size();
    std::printf("Q %d %d %d %d %zu\n", q[0], q[1], q[2], q[3], n);
    for (size_t i = 0; i < n; ++i) std::printf("%.17g\n", buf[0][i]);
  }
  // batched extension gives the same numbers
  std::vector<std::array<int, 4>> qs;
  for (auto& q : quartets) qs.push_back({q[0], q[1], q[2], q[3]});
  std::vector<int64_t> off;
  auto all = e.compute_batch(basis, qs, &off);
  e.compute(basis[2], basis[3], basis[2], basis[3]);
  for (size_t i = 0; i < 5 * 6 * 5 * 6; ++i) CHECK(all[off[4] + i] == buf[0][i]);
  // 3- and 2-centre brakets: the missing centre is Shell::unit() (printed
  // as shell 4 for the oracle)
  const double eps = std::numeric_limits<double>::epsilon();
  Engine e3(Operator::coulomb, 2, 2, 0, eps, 0, BraKet::xs_xx);
  e3.compute(basis[2], basis[0], basis[1]);
  CHECK(e3.results()[0] != nullptr);
  {
    const size_t n = basis[2].size() * basis[0].size() * basis[1].size();
    std::printf("Q 2 4 0 1 %zu\n", n);
    for (size_t i = 0; i < n; ++i) std::printf("%.17g\n", e3.results()[0][i]);
  }
  Engine e2(Operator::coulomb, 2, 2, 0, eps, 0, BraKet::xs_xs);
  e2.compute(basis[1], basis[2]);
  CHECK(e2.results()[0] != nullptr);
  {
    const size_t n = basis[1].size() * basis[2].size();
    std::printf("Q 1 4 2 4 %zu\n", n);
    for (size_t i = 0; i < n; ++i) std::printf("%.17g\n", e2.results()[0][i]);
  }
  // erf-attenuated Coulomb, omega = 0.7 as the Engine's params (printed as R)
  Engine er(Operator::erf_coulomb, 2, 2, 0, eps, 0.7);
  er.compute(basis[0], basis[1], basis[2], basis[3]);
  CHECK(er.results()[0] != nullptr);
  {
    const size_t n = basis[0].size() * basis[1].size() * basis[2].size() * basis[3].size();
    std::printf("R 0 1 2 3 %zu\n", n);
    for (size_t i = 0; i < n; ++i) std::printf("%.17g\n", er.results()[0][i]);
  }
  // first derivatives (deriv_order = 1): 12 shell sets, printed as D
  Engine ed(Operator::coulomb, 2, 2, 1);
  CHECK(ed.nshellsets() == 12);
  CHECK(throws<Engine::lmax_exceeded>([] { Engine(Operator::coulomb, 3, l2g_max_l(), 1); }));
  ed.compute(basis[0], basis[1], basis[2], basis[3]);
  {
    const size_t n = basis[0].size() * basis[1].size() * basis[2].size() * basis[3].size();
    std::printf("D 0 1 2 3 %zu\n", n);
    for (int dset = 0; dset < 12; ++dset) {
      CHECK(ed.results()[dset] != nullptr);
      for (size_t i = 0; i < n; ++i) std::printf("%.17g\n", ed.results()[dset][i]);
    }
  }
  // one-body operators as compute_1body_ints drives them (hartree-fock.cc:480-535),
  // printed as O (overlap), K (kinetic), N (nuclear with two charges)
  {
    Engine eo(Operator::overlap, 2, 2);
    CHECK(eo.braket() == BraKet::x_x);
    eo.compute(basis[2], basis[1]);
    const size_t n = basis[2].size() * basis[1].size();
    std::printf("O 2 1 %zu\n", n);
    for (size_t i = 0; i < n; ++i) std::printf("%.17g\n", eo.results()[0][i]);
    Engine ek(Operator::kinetic, 2, 2);
    ek.compute(basis[3], basis[2]);
    const size_t nk = basis[3].size() * basis[2].size();
    std::printf("K 3 2 %zu\n", nk);
    for (size_t i = 0; i < nk; ++i) std::printf("%.17g\n", ek.results()[0][i]);
    Engine en(Operator::nuclear, 2, 2);
    CHECK(throws<std::invalid_argument>([&] { en.compute(basis[1], basis[3]); }));  // no charges yet
    std::vector<std::pair<double, std::array<double, 3>>> q{{8.0, {{0.1, 0.0, -0.2}}}, {1.0, {{1.0, 0.5, 0.3}}}};
    en.set_params(q);
    en.compute(basis[1], basis[3]);
    const size_t nn = basis[1].size() * basis[3].size();
    std::printf("N 1 3 %zu\n", nn);
    for (size_t i = 0; i < nn; ++i) std::printf("%.17g\n", en.results()[0][i]);
    // prescaled overlap with uniform Cartesian normalization, printed as U
    // (Cartesian d against p)
    Engine eu(Operator::overlap, 2, 2);
    eu.prescale_by(1.7).set(libint2::CartesianShellNormalization::uniform);
    eu.compute(basis[3], basis[1]);
    const size_t nu = basis[3].size() * basis[1].size();
    std::printf("U 3 1 %zu\n", nu);
    for (size_t i = 0; i < nu; ++i) std::printf("%.17g\n", eu.results()[0][i]);
  }
  // python/tests/test_libint2.py:39-42: ||(s s|s)|| = 3.6563211198 for
  // s = Shell(0, [(1, 10)])
  Shell su{{1.0}, {{0, false, {10.0}}}, {{0, 0, 0}}};
  Engine ex(Operator::coulomb, 1, 0, 0, 1e-15, 0, BraKet::xs_xx);
  ex.compute(su, su, su);
  CHECK(ex.results()[0] != nullptr && std::abs(ex.results()[0][0] - 3.6563211198) < 1e-9);
  // fully screened quartet -> nullptr (engine.impl.h:1781-1784)
  Shell t1{{80.0}, {{0, false, {1.0}}}, {{0, 0, 0}}};
  Shell t2{{80.0}, {{0, false, {1.0}}}, {{6.0, 0, 0}}};
  e.set_precision(1e-10);
  e.compute(t1, t2, t1, t2);
  CHECK(buf[0] == nullptr);
  // compute2 with precomputed shell pairs: the pairs' screening method must
  // be the Engine's; a Schwarz method needs the pairs
  {
    // pairs screened at least as finely as the Engine's precision are used
    // as given (a coarser pair would be rebuilt, engine.impl.h:1259-1276)
    Engine es(Operator::coulomb, 2, 2);
    ShellPair p01(basis[0], basis[1], std::log(1e-20), ScreeningMethod::Original);
    ShellPair p23(basis[2], basis[3], std::log(1e-20), ScreeningMethod::Original);
    es.compute2<Operator::coulomb, BraKet::xx_xx, 0>(basis[0], basis[1], basis[2], basis[3], &p01, &p23);
    CHECK(es.results()[0] != nullptr);
    es.set(ScreeningMethod::Conservative);
    CHECK(throws<std::invalid_argument>([&] {
      es.compute2<Operator::coulomb, BraKet::xx_xx, 0>(basis[0], basis[1], basis[2], basis[3], &p01, &p23);
    }));
    es.set(ScreeningMethod::SchwarzInf);
    CHECK(throws<std::invalid_argument>([&] { es.compute(basis[0], basis[1], basis[2], basis[3]); }));
    CHECK(throws<std::invalid_argument>([&] {
      es.compute2<Operator::coulomb, BraKet::xx_xx, 0>(basis[0], basis[1], basis[2], basis[3], &p01, nullptr);
    }));
    // the dispatch table: the coulomb / xx_xx / 0 entry is compute2 itself
    es.set(ScreeningMethod::Original);
    auto ptr = es.compute2_ptrs()[((size_t)((int)Operator::coulomb - (int)Operator::first_2body_oper) *
                                       libint2::nbrakets_2body +
                                   (size_t)((int)BraKet::xx_xx - (int)BraKet::first_2body_braket)) *
                                      libint2::nderivorders_2body];
    CHECK(ptr != nullptr);
    (es.*ptr)(basis[2], basis[3], basis[2], basis[3], nullptr, nullptr);
    for (size_t i = 0; i < 5 * 6 * 5 * 6; ++i) CHECK(es.results()[0][i] == all[off[4] + i]);
  }
  // the Fock loop of hartree-fock++ through compute2 with SchwarzInf pairs,
  // printed as F nbf, then D and G (row-major)
  {
    const auto obs = fock_basis();
    size_t nbf = 0;
    for (auto& x : obs) nbf += x.size();
    std::vector<double> C(nbf * 4), D(nbf * nbf, 0.0);
    unsigned long long st = 12345;
    for (auto& c : C) {
      st = st * 6364136223846793005ULL + 1442695040888963407ULL;
      c = 0.2 * ((double)(st >> 11) / 9007199254740992.0 - 0.5);
    }
    for (size_t i = 0; i < nbf; ++i)
      for (size_t j = 0; j < nbf; ++j)
        for (size_t k = 0; k < 4; ++k) D[i * nbf + j] += C[i * 4 + k] * C[j * 4 + k];
    const auto G = fock_loop(obs, D, nbf, 1e-12);
    std::printf("F %zu\n", nbf);
    for (double x : D) std::printf("%.17g\n", x);
    for (double x : G) std::printf("%.17g\n", x);
  }
  // per-call latency of Engine::compute on the resident evaluator: a
  // contracted (dd|dd) quartet, repeated
  {
    Shell dd{{2.2, 0.6, 0.2}, {{2, true, {0.3, 0.6, 0.4}}}, {{0.1, -0.2, 0.3}}};
    Shell de{{1.5, 0.4}, {{2, true, {0.5, 0.7}}}, {{1.1, 0.4, -0.6}}};
    Engine el(Operator::coulomb, 3, 2);
    el.compute(dd, de, dd, de);
    const int reps = 2000;
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) el.compute(dd, de, dd, de);
    const double us =
        std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / reps;
    std::printf("LAT %.3f\n", us);
  }
  libint2::finalize();
  std::fprintf(stderr, "facade checks: %s\n", fails ? "FAIL" : "ok");
  return fails ? 1 : 0;
}
