// Drives the libint2-compatible façade (include/libint2_amd/engine.h) the way
// libint's own tests drive Engine (tests/unit/test-core.cc:59-77,
// test-2body.cc): prints every integral of a few quartets for comparison
// with the oracle, and checks the error behaviour.
//   facade_test cpu   -> only the checks that need no GPU
//   facade_test       -> everything (needs a GPU)
#include <cstdio>
#include <cstring>
#include <limits>
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
  libint2::finalize();
  std::fprintf(stderr, "facade checks: %s\n", fails ? "FAIL" : "ok");
  return fails ? 1 : 0;
}
