// libint2_amd/engine.h — header-only C++ façade that keeps libint's
// Engine(Operator::coulomb, max_nprim, max_l) / compute(s1, s2, s3, s4)
// surface (include/libint2/engine.h:444-1064) and evaluates on an MI355X
// through the C ABI of l2g.h.  Include this instead of <libint2.hpp> for the
// two-electron Coulomb path; link with -ll2g.
//
// Kept from the reference, with the same meaning:
//   * Shell with renormalised coefficients (shell.h:720-1012); Shell::unit()
//   * Operator::coulomb, erf_coulomb and erfc_coulomb (params = omega),
//     BraKet::xx_xx and the 3-/2-centre xs_xx, xx_xs, xs_xs
//     (compute with 3 or 2 shells, the missing centre Shell::unit();
//     engine.impl.h:140-178), ScreeningMethod (shell.h:1041-1059),
//     CartesianShellNormalization, default_screening_method()
//   * initialize() / finalize() / initialized() (initialize.h:76-113),
//     set_solid_harmonics_ordering (:117-119)
//   * Engine ctor (engine.h:502-526), compute (:747-749), compute2 (:787-791),
//     results() (:731), set_precision (:809-823), set(ScreeningMethod),
//     set(CartesianShellNormalization), prescale_by (:857), max_nprim(),
//     max_l(), set_max_nprim, set_max_l; exceptions lmax_exceeded
//     (:893-916) and using_default_initialized (:919-926)
//   * results()[0] == nullptr when every primitive was screened out
//     (engine.impl.h:1781-1784); buffer valid until the next compute call.
// Added: Engine::compute_batch for many quartets in one GPU launch set.
#pragma once

#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <type_traits>

#include "../l2g.h"

namespace libint2 {

// ---------------------------------------------------------------- enums
enum class Operator { overlap = 0, kinetic, nuclear, coulomb = 7, erf_coulomb, erfc_coulomb, invalid = -1 };
enum class BraKet { x_x = 0, xx_xx = 1, xs_xx, xx_xs, xs_xs, invalid = -1 };
enum class ScreeningMethod {
  Original = L2G_SCREEN_ORIGINAL,
  Conservative = L2G_SCREEN_CONSERVATIVE,
  Schwarz = L2G_SCREEN_SCHWARZ,
  SchwarzInf = L2G_SCREEN_SCHWARZINF,
  Invalid = 0
};
enum class CartesianShellNormalization { standard, uniform };
enum SHGShellOrdering { SHGShellOrdering_Standard = L2G_SHG_STANDARD, SHGShellOrdering_Gaussian = L2G_SHG_GAUSSIAN };

namespace detail {
inline ScreeningMethod& default_screening_method_ref() {
  static ScreeningMethod m = ScreeningMethod::Original;
  return m;
}
struct Runtime {
  l2g_context ctx = nullptr;
  int ordering = SHGShellOrdering_Standard;
  std::mutex mtx;
  static Runtime& get() {
    static Runtime r;
    return r;
  }
};
inline void check(int rc) {
  if (rc == 0) return;
  const std::string msg = l2g_last_error();
  if (rc == L2G_E_INVALID) throw std::invalid_argument(msg);
  throw std::runtime_error("l2g error " + std::to_string(rc) + ": " + msg);
}
}  // namespace detail

inline ScreeningMethod default_screening_method() { return detail::default_screening_method_ref(); }
inline void default_screening_method(ScreeningMethod m) { detail::default_screening_method_ref() = m; }

inline void set_solid_harmonics_ordering(SHGShellOrdering o) { detail::Runtime::get().ordering = o; }
inline SHGShellOrdering solid_harmonics_ordering() { return (SHGShellOrdering)detail::Runtime::get().ordering; }

inline bool initialized() { return detail::Runtime::get().ctx != nullptr; }

/// opens the GPU context (device 0 unless given)
inline void initialize(int device = 0) {
  auto& r = detail::Runtime::get();
  std::lock_guard<std::mutex> lk(r.mtx);
  if (r.ctx) return;
  detail::check(l2g_init(device, &r.ctx));
  detail::check(l2g_set_solid_harmonics_ordering(r.ctx, r.ordering));
}
inline void finalize() {
  auto& r = detail::Runtime::get();
  std::lock_guard<std::mutex> lk(r.mtx);
  if (!r.ctx) return;
  l2g_finalize(r.ctx);
  r.ctx = nullptr;
}

// ---------------------------------------------------------------- Shell
struct Shell {
  typedef double real_t;
  struct Contraction {
    int l;
    bool pure;
    std::vector<real_t> coeff;
    size_t cartesian_size() const { return (l + 1) * (l + 2) / 2; }
    size_t size() const { return pure ? (2 * l + 1) : cartesian_size(); }
  };
  std::vector<real_t> alpha;
  std::vector<Contraction> contr;
  std::array<real_t, 3> O{{0, 0, 0}};
  std::vector<real_t> max_ln_coeff;

  Shell() = default;
  /// @param embed_normalization_into_coefficients as libint's Shell
  Shell(std::vector<real_t> a, std::vector<Contraction> c, std::array<real_t, 3> origin,
        bool embed_normalization_into_coefficients = true)
      : alpha(std::move(a)), contr(std::move(c)), O(origin) {
    max_ln_coeff.assign(alpha.size(), 0.0);
    for (auto& cc : contr) {
      if (cc.coeff.size() != alpha.size()) throw std::invalid_argument("Shell: coefficient/exponent count mismatch");
      if (embed_normalization_into_coefficients)
        detail::check(l2g_shell_renorm(cc.l, (int)alpha.size(), alpha.data(), cc.coeff.data(), nullptr,
                                       do_enforce_unit_normalization() ? 1 : 0));
    }
    for (size_t p = 0; p < alpha.size(); ++p) {
      real_t m = -std::numeric_limits<real_t>::max();
      for (auto& cc : contr) m = std::max(m, std::log(std::abs(cc.coeff[p])));
      max_ln_coeff[p] = m;
    }
  }
  size_t nprim() const { return alpha.size(); }
  size_t ncontr() const { return contr.size(); }
  size_t size() const {
    size_t s = 0;
    for (auto& c : contr) s += c.size();
    return s;
  }
  size_t cartesian_size() const {
    size_t s = 0;
    for (auto& c : contr) s += c.cartesian_size();
    return s;
  }
  static bool do_enforce_unit_normalization(int flag = -1) {
    static bool v = true;
    if (flag >= 0) v = flag != 0;
    return v;
  }
  /// exponent 0, coefficient 1, at the origin (shell.h:904-909)
  static const Shell& unit() {
    static const Shell u = [] {
      Shell s;
      s.alpha = {0.0};
      s.contr = {Contraction{0, false, {1.0}}};
      s.max_ln_coeff = {0.0};
      return s;
    }();
    return u;
  }
};

// ---------------------------------------------------------------- Engine
class Engine {
 public:
  typedef double scalar_type;
  typedef std::vector<const scalar_type*> target_ptr_vec;

  struct lmax_exceeded : public std::logic_error {
    lmax_exceeded(const char* task, size_t lmax_limit, size_t lmax_requested)
        : std::logic_error("Engine::lmax_exceeded -- angular momentum limit exceeded"),
          task_(task),
          lmax_limit_(lmax_limit),
          lmax_requested_(lmax_requested) {}
    const char* task() const { return task_.c_str(); }
    size_t lmax_limit() const { return lmax_limit_; }
    size_t lmax_requested() const { return lmax_requested_; }

   private:
    std::string task_;
    size_t lmax_limit_, lmax_requested_;
  };
  struct using_default_initialized : public std::logic_error {
    using_default_initialized() : std::logic_error("Engine::using_default_initialized -- attempt to use a default-initialized Engine") {}
  };

  Engine() = default;
  /// params: omega (the attenuation) for erf_coulomb / erfc_coulomb, as
  /// operator_traits<erf_coulomb>::oper_params_type; ignored for coulomb
  template <typename Params = int>
  Engine(Operator oper, size_t max_nprim, int max_l, int deriv_order = 0,
         scalar_type precision = std::numeric_limits<scalar_type>::epsilon(), Params params = Params(),
         BraKet braket = BraKet::invalid, ScreeningMethod screening_method = default_screening_method())
      : oper_(oper),
        braket_(braket == BraKet::invalid ? (is_onebody(oper) ? BraKet::x_x : BraKet::xx_xx) : braket),
        screening_(screening_method) {
    assert(max_nprim > 0);
    check_operator(oper);
    set_params(params);
    if ((braket_ == BraKet::x_x) != is_onebody(oper))
      throw std::invalid_argument("libint2_amd: BraKet::x_x goes with overlap, kinetic and nuclear only");
    if (is_onebody(oper) && deriv_order != 0)
      throw std::invalid_argument("libint2_amd: one-body integrals are provided for deriv_order 0");
    if (deriv_order < 0 || deriv_order > 1) throw std::invalid_argument("libint2_amd: deriv_order 0 or 1 is provided");
    if (deriv_order == 1 && (oper != Operator::coulomb || braket_ != BraKet::xx_xx))
      throw std::invalid_argument("libint2_amd: deriv_order 1 is provided for 4-centre Coulomb integrals");
    assert(initialized() && "libint2::initialize() must be called before constructing an Engine");
    // first derivatives evaluate ERIs with every shell raised by one
    if (max_l + deriv_order > l2g_max_l()) throw lmax_exceeded("coulomb", l2g_max_l() - deriv_order, max_l);
    deriv_order_ = deriv_order;
    max_nprim_ = max_nprim;
    lmax_ = max_l;
    set_precision(precision);
    targets_.assign(nshellsets(), nullptr);
    default_ = false;
  }

  Engine& set_precision(scalar_type prec) {
    precision_ = prec <= 0 ? 0.0 : prec;
    return *this;
  }
  scalar_type precision() const { return precision_; }
  Engine& set(ScreeningMethod m) {
    screening_ = m;
    return *this;
  }
  Engine& set(CartesianShellNormalization n) {
    uniform_ = (n == CartesianShellNormalization::uniform);
    return *this;
  }
  Engine& set(Operator o) {
    if (default_) throw using_default_initialized();
    check_operator(o);
    if (is_onebody(o) != is_onebody(oper_))
      throw std::invalid_argument("libint2_amd: set(Operator) cannot switch between one- and two-body operators");
    if (deriv_order_ == 1 && o != Operator::coulomb)
      throw std::invalid_argument("libint2_amd: deriv_order 1 is provided for Coulomb integrals");
    oper_ = o;
    return *this;
  }
  /// omega of erf_coulomb / erfc_coulomb (engine.h:685-690)
  /// or the point charges {q, {x, y, z}} of Operator::nuclear (engine.h:685-690)
  template <typename Params>
  Engine& set_params(const Params& params) {
    if constexpr (std::is_arithmetic_v<Params>) omega_ = (double)params;
    if constexpr (std::is_convertible_v<Params, std::vector<std::pair<double, std::array<double, 3>>>>) {
      charges_.clear();
      for (const auto& c : params) charges_.push_back(c);
    }
    return *this;
  }
  Engine& set(BraKet b) {
    if (default_) throw using_default_initialized();
    if (b == BraKet::invalid || (b == BraKet::x_x) != is_onebody(oper_))
      throw std::invalid_argument("libint2_amd: BraKet::x_x goes with overlap, kinetic and nuclear only");
    if (deriv_order_ == 1 && b != BraKet::xx_xx)
      throw std::invalid_argument("libint2_amd: deriv_order 1 is provided for 4-centre integrals");
    braket_ = b;
    return *this;
  }
  BraKet braket() const { return braket_; }
  Engine& prescale_by(scalar_type s) {
    scale_ = s;
    return *this;
  }
  scalar_type prescaled_by() const { return scale_; }
  Engine& set_max_nprim(size_t n) {
    if (default_) throw using_default_initialized();
    max_nprim_ = n;
    return *this;
  }
  Engine& set_max_l(int l) {
    if (default_) throw using_default_initialized();
    if (l + deriv_order_ > l2g_max_l()) throw lmax_exceeded("coulomb", l2g_max_l() - deriv_order_, l);
    lmax_ = l;
    return *this;
  }
  size_t max_nprim() const { return max_nprim_; }
  int max_l() const { return lmax_; }
  /// 1, or for deriv_order 1 the 12 derivative shell sets d/dA_x .. d/dD_z
  /// (CartesianDerivIterator<4> order, util/intpart_iter.h:89-103)
  unsigned int nshellsets() const { return deriv_order_ == 1 ? 12 : 1; }
  int deriv_order() const { return deriv_order_; }
  const target_ptr_vec& results() const { return targets_; }

  /// one shell quartet, caller order (s1 s2|s3 s4), libint's layout
  const target_ptr_vec& compute(const Shell& s1, const Shell& s2, const Shell& s3, const Shell& s4) {
    if (default_) throw using_default_initialized();
    assert(braket_ == BraKet::xx_xx && "# of arguments to compute() does not match the braket type");
    return compute_quartet(s1, s2, s3, s4);
  }
  /// 3-centre: (a|cd) for BraKet::xs_xx, (ab|c) for BraKet::xx_xs; the
  /// missing centre is Shell::unit(), as Engine::compute (engine.impl.h:140-178)
  const target_ptr_vec& compute(const Shell& s1, const Shell& s2, const Shell& s3) {
    if (default_) throw using_default_initialized();
    assert((braket_ == BraKet::xs_xx || braket_ == BraKet::xx_xs) &&
           "# of arguments to compute() does not match the braket type");
    return braket_ == BraKet::xs_xx ? compute_quartet(s1, Shell::unit(), s2, s3)
                                    : compute_quartet(s1, s2, s3, Shell::unit());
  }
  /// 2-centre: (a|c) for BraKet::xs_xs; <a|O|b> for the one-body operators
  /// (BraKet::x_x): overlap, kinetic, nuclear (charges from set_params)
  const target_ptr_vec& compute(const Shell& s1, const Shell& s2) {
    if (default_) throw using_default_initialized();
    if (braket_ == BraKet::x_x) return compute_onebody(s1, s2);
    assert(braket_ == BraKet::xs_xs && "# of arguments to compute() does not match the braket type");
    return compute_quartet(s1, Shell::unit(), s2, Shell::unit());
  }

 private:
  const target_ptr_vec& compute_quartet(const Shell& s1, const Shell& s2, const Shell& s3, const Shell& s4) {
    const Shell* sh[4] = {&s1, &s2, &s3, &s4};
    for (auto* s : sh) {
      assert(s->ncontr() == 1 && "generally-contracted shells are not yet supported");
      if (s->contr[0].l > lmax_) throw lmax_exceeded("coulomb", lmax_, s->contr[0].l);
    }
    std::vector<int> l(4), pure(4), np(4);
    std::vector<double> a, c, o;
    for (int k = 0; k < 4; ++k) {
      l[k] = sh[k]->contr[0].l;
      pure[k] = sh[k]->contr[0].pure ? 1 : 0;
      np[k] = (int)sh[k]->nprim();
      a.insert(a.end(), sh[k]->alpha.begin(), sh[k]->alpha.end());
      c.insert(c.end(), sh[k]->contr[0].coeff.begin(), sh[k]->contr[0].coeff.end());
      o.insert(o.end(), sh[k]->O.begin(), sh[k]->O.end());
    }
    auto& rt = detail::Runtime::get();
    detail::check(l2g_set_solid_harmonics_ordering(rt.ctx, rt.ordering));
    l2g_basis b = nullptr;
    detail::check(l2g_basis_create(rt.ctx, 4, l.data(), pure.data(), np.data(), a.data(), c.data(), o.data(), &b));
    std::unique_ptr<l2g_basis_s, int (*)(l2g_basis)> hold(b, &l2g_basis_destroy);
    const int32_t q[4] = {0, 1, 2, 3};
    if (deriv_order_ == 1) {
      if (uniform_ || scale_ != 1.0)
        throw std::invalid_argument("libint2_amd: deriv_order 1 with prescaling or uniform normalization is not provided");
      l2g_deriv1_plan dp = nullptr;
      detail::check(l2g_eri_deriv1_plan_create(b, 1, q, precision_, method_for_plan(), &dp));
      std::unique_ptr<l2g_deriv1_plan_s, int (*)(l2g_deriv1_plan)> hd(dp, &l2g_eri_deriv1_plan_destroy);
      int64_t n = 0;
      detail::check(l2g_eri_deriv1_plan_size(dp, &n));
      buffer_.resize((size_t)std::max<int64_t>(12 * n, 1));
      detail::check(l2g_eri_deriv1_plan_execute(dp, nullptr));
      int8_t nz = 0;
      detail::check(l2g_eri_deriv1_plan_download(dp, buffer_.data(), &nz));
      for (int d = 0; d < 12; ++d) targets_[d] = nz ? buffer_.data() + d * n : nullptr;
      return targets_;
    }
    l2g_eri_plan p = nullptr;
    detail::check(l2g_eri_plan_create(b, 1, q, precision_, method_for_plan(), scale_, uniform_ ? 1 : 0, &p));
    std::unique_ptr<l2g_eri_plan_s, int (*)(l2g_eri_plan)> hp(p, &l2g_eri_plan_destroy);
    detail::check(l2g_eri_plan_set_operator(p, plan_operator(), omega_));
    int64_t n = 0;
    detail::check(l2g_eri_plan_size(p, &n));
    buffer_.resize((size_t)std::max<int64_t>(n, 1));
    detail::check(l2g_eri_plan_execute(p, nullptr));
    int8_t nz = 0;
    detail::check(l2g_eri_plan_download(p, buffer_.data(), &nz));
    targets_[0] = nz ? buffer_.data() : nullptr;
    return targets_;
  }


 public:
  template <Operator op, BraKet bk, size_t der>
  const target_ptr_vec& compute2(const Shell& b1, const Shell& b2, const Shell& k1, const Shell& k2,
                                 const void* spbra = nullptr, const void* spket = nullptr) {
    static_assert(op == Operator::coulomb && bk == BraKet::xx_xx && der <= 1,
                  "libint2_amd provides compute2<coulomb, xx_xx, 0 or 1> only");
    if ((int)der != deriv_order_) throw std::invalid_argument("libint2_amd: compute2's deriv order differs from the Engine's");
    assert(((spbra == nullptr) == (spket == nullptr)) && "Engine::compute2 -- expects zero or two ShellPair objects");
    return compute_quartet(b1, b2, k1, k2);
  }

  /// Batched extension: every quartet (indices into `basis`, caller order)
  /// in one set of GPU launches; returns the flat results and fills
  /// offsets[q] (in doubles) and nonzero[q].
  std::vector<double> compute_batch(const std::vector<Shell>& basis, const std::vector<std::array<int, 4>>& quartets,
                                    std::vector<int64_t>* offsets = nullptr, std::vector<int8_t>* nonzero = nullptr) {
    if (default_) throw using_default_initialized();
    std::vector<int> l, pure, np;
    std::vector<double> a, c, o;
    for (auto& s : basis) {
      if (s.contr[0].l > lmax_) throw lmax_exceeded("coulomb", lmax_, s.contr[0].l);
      l.push_back(s.contr[0].l);
      pure.push_back(s.contr[0].pure);
      np.push_back((int)s.nprim());
      a.insert(a.end(), s.alpha.begin(), s.alpha.end());
      c.insert(c.end(), s.contr[0].coeff.begin(), s.contr[0].coeff.end());
      o.insert(o.end(), s.O.begin(), s.O.end());
    }
    auto& rt = detail::Runtime::get();
    detail::check(l2g_set_solid_harmonics_ordering(rt.ctx, rt.ordering));
    l2g_basis b = nullptr;
    detail::check(l2g_basis_create(rt.ctx, (int)basis.size(), l.data(), pure.data(), np.data(), a.data(), c.data(),
                                   o.data(), &b));
    std::unique_ptr<l2g_basis_s, int (*)(l2g_basis)> hold(b, &l2g_basis_destroy);
    std::vector<int32_t> q;
    for (auto& x : quartets) q.insert(q.end(), x.begin(), x.end());
    l2g_eri_plan p = nullptr;
    detail::check(l2g_eri_plan_create(b, (int64_t)quartets.size(), q.data(), precision_, method_for_plan(), scale_,
                                      uniform_ ? 1 : 0, &p));
    std::unique_ptr<l2g_eri_plan_s, int (*)(l2g_eri_plan)> hp(p, &l2g_eri_plan_destroy);
    detail::check(l2g_eri_plan_set_operator(p, plan_operator(), omega_));
    int64_t n = 0;
    detail::check(l2g_eri_plan_size(p, &n));
    std::vector<double> out((size_t)std::max<int64_t>(n, 1));
    if (offsets) {
      offsets->resize(quartets.size());
      detail::check(l2g_eri_plan_offsets(p, offsets->data()));
    }
    detail::check(l2g_eri_plan_execute(p, nullptr));
    std::vector<int8_t> nz(quartets.size());
    detail::check(l2g_eri_plan_download(p, out.data(), nz.data()));
    if (nonzero) *nonzero = nz;
    out.resize((size_t)n);
    return out;
  }

 private:
  int method_for_plan() const {
    // without precomputed shell pairs libint allows Original/Conservative only
    // (engine.impl.h:1212-1216)
    if (screening_ != ScreeningMethod::Original && screening_ != ScreeningMethod::Conservative)
      throw std::invalid_argument(
          "Engine::compute2: without precomputed shell pair data can only use original or conservative screening methods");
    return (int)screening_;
  }

  static bool is_onebody(Operator o) {
    return o == Operator::overlap || o == Operator::kinetic || o == Operator::nuclear;
  }
  static void check_operator(Operator o) {
    if (o != Operator::coulomb && o != Operator::erf_coulomb && o != Operator::erfc_coulomb && !is_onebody(o))
      throw std::invalid_argument(
          "libint2_amd: Operator::overlap, kinetic, nuclear, coulomb, erf_coulomb and erfc_coulomb are provided");
  }
  // <s1|O|s2> as the (s1, s2) block of the two-shell basis's matrix, row-major
  const target_ptr_vec& compute_onebody(const Shell& s1, const Shell& s2) {
    const Shell* sh[2] = {&s1, &s2};
    std::vector<int> l(2), pure(2), np(2);
    std::vector<double> a, c, o;
    for (int k = 0; k < 2; ++k) {
      assert(sh[k]->ncontr() == 1 && "generally-contracted shells are not yet supported");
      if (sh[k]->contr[0].l > lmax_) throw lmax_exceeded("1-body", lmax_, sh[k]->contr[0].l);
      l[k] = sh[k]->contr[0].l;
      pure[k] = sh[k]->contr[0].pure ? 1 : 0;
      np[k] = (int)sh[k]->nprim();
      a.insert(a.end(), sh[k]->alpha.begin(), sh[k]->alpha.end());
      c.insert(c.end(), sh[k]->contr[0].coeff.begin(), sh[k]->contr[0].coeff.end());
      o.insert(o.end(), sh[k]->O.begin(), sh[k]->O.end());
    }
    if (oper_ == Operator::nuclear && charges_.empty())
      throw std::invalid_argument("libint2_amd: Operator::nuclear needs the charges (set_params)");
    auto& rt = detail::Runtime::get();
    detail::check(l2g_set_solid_harmonics_ordering(rt.ctx, rt.ordering));
    l2g_basis b = nullptr;
    detail::check(l2g_basis_create(rt.ctx, 2, l.data(), pure.data(), np.data(), a.data(), c.data(), o.data(), &b));
    std::unique_ptr<l2g_basis_s, int (*)(l2g_basis)> hold(b, &l2g_basis_destroy);
    const size_t n1 = s1.size(), n2 = s2.size(), n = n1 + n2;
    std::vector<double> q, xyz, M(n * n);
    for (const auto& ch : charges_) {
      q.push_back(ch.first);
      xyz.insert(xyz.end(), ch.second.begin(), ch.second.end());
    }
    double* S = oper_ == Operator::overlap ? M.data() : nullptr;
    double* T = oper_ == Operator::kinetic ? M.data() : nullptr;
    double* V = oper_ == Operator::nuclear ? M.data() : nullptr;
    detail::check(l2g_onebody(b, V ? (int)q.size() : 0, V ? q.data() : nullptr, V ? xyz.data() : nullptr, S, T, V));
    buffer_.resize(std::max<size_t>(n1 * n2, 1));
    for (size_t i = 0; i < n1; ++i)
      for (size_t j = 0; j < n2; ++j) buffer_[i * n2 + j] = M[i * n + n1 + j];
    targets_[0] = buffer_.data();
    return targets_;
  }
  int plan_operator() const {
    return oper_ == Operator::erf_coulomb ? L2G_OPER_ERF_COULOMB
                                          : (oper_ == Operator::erfc_coulomb ? L2G_OPER_ERFC_COULOMB : L2G_OPER_COULOMB);
  }

  Operator oper_ = Operator::invalid;
  double omega_ = 0;
  std::vector<std::pair<double, std::array<double, 3>>> charges_;  // Operator::nuclear
  int deriv_order_ = 0;
  BraKet braket_ = BraKet::invalid;
  ScreeningMethod screening_ = ScreeningMethod::Original;
  size_t max_nprim_ = 0;
  int lmax_ = -1;
  scalar_type precision_ = std::numeric_limits<scalar_type>::epsilon();
  scalar_type scale_ = 1;
  bool uniform_ = false;
  bool default_ = true;
  target_ptr_vec targets_;
  std::vector<double> buffer_;
};

}  // namespace libint2
