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
// libint's operator list and values (engine.h:100-185); this build evaluates
// overlap, kinetic, nuclear, coulomb, erf_coulomb and erfc_coulomb, and an
// Engine for any other operator throws std::invalid_argument
enum class Operator {
  overlap = 0,
  kinetic,
  nuclear,
  erf_nuclear,
  erfc_nuclear,
  emultipole1,
  emultipole2,
  emultipole3,
  sphemultipole,
  opVop,
  delta,
  coulomb,
  r12_m1 = coulomb,
  cgtg,
  cgtg_x_coulomb,
  delcgtg2,
  r12,
  r12_1 = r12,
  erf_coulomb,
  erfc_coulomb,
  stg,
  stg_x_coulomb,
  yukawa = stg_x_coulomb,
  invalid = -1,
  first_1body_oper = overlap,
  last_1body_oper = opVop,
  first_2body_oper = delta,
  last_2body_oper = stg_x_coulomb,
  first_oper = first_1body_oper,
  last_oper = last_2body_oper
};
// braket.h:36-49
enum class BraKet {
  x_x = 0,
  xx_xx,
  xs_xx,
  xx_xs,
  xs_xs,
  invalid = -1,
  first_1body_braket = x_x,
  last_1body_braket = x_x,
  first_2body_braket = xx_xx,
  last_2body_braket = xs_xs,
  first_braket = first_1body_braket,
  last_braket = last_2body_braket
};
// sizes of Engine::compute2_ptrs() (engine.h:431-437); derivatives to first order
constexpr size_t nopers_2body = static_cast<int>(Operator::last_2body_oper) -
                                static_cast<int>(Operator::first_2body_oper) + 1;
constexpr size_t nbrakets_2body = static_cast<int>(BraKet::last_2body_braket) -
                                  static_cast<int>(BraKet::first_2body_braket) + 1;
constexpr size_t nderivorders_2body = 2;
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
  /// primitive p as a shell of its own with coefficient 1, renormalised only
  /// if unit_normalized (shell.h:936-944)
  Shell extract_primitive(size_t p, bool unit_normalized = true) const {
    std::vector<Contraction> c;
    for (auto& x : contr) c.push_back(Contraction{x.l, x.pure, {1.0}});
    return Shell({alpha.at(p)}, c, O, unit_normalized);
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

// ---------------------------------------------------------------- ShellPair
// Primitive-pair data of an ordered shell pair, the type libint's drivers
// precompute once per significant pair and hand to Engine::compute2
// (shell.h:1079-1329): which primitive pairs survive the screen and the
// logarithm of each one's screening factor.  The integrals themselves do not
// read P, K and 1/gamma from here (the device path recomputes them from the
// shells, with the same formulas); they are kept for callers that inspect
// the pair.
struct ShellPair {
  typedef Shell::real_t real_t;
  struct PrimPairData {
    real_t P[3];               // (a1 A + a2 B) / (a1 + a2)
    real_t K;                  // sqrt(2) pi^(5/4) exp(-a1 a2 / (a1 + a2) |AB|^2) / (a1 + a2)
    real_t one_over_gamma;     // 1 / (a1 + a2)
    real_t nonsph_screen_fac;  // Conservative screening only
    real_t ln_scr;             // ln of the primitive pair's screening factor
    int p1;
    int p2;
  };
  std::vector<PrimPairData> primpairs;
  real_t AB[3] = {0, 0, 0};
  real_t ln_prec = std::numeric_limits<real_t>::lowest();
  ScreeningMethod screening_method_ = ScreeningMethod::Invalid;

  ShellPair() = default;
  explicit ShellPair(size_t max_nprim) { primpairs.reserve(max_nprim * max_nprim); }
  template <typename Real>
  ShellPair(const Shell& s1, const Shell& s2, Real ln_prec_,
            ScreeningMethod method = default_screening_method()) {
    init(s1, s2, ln_prec_, method);
  }
  template <typename Real, typename SchwarzFactorEvaluator>
  ShellPair(const Shell& s1, const Shell& s2, Real ln_prec_, ScreeningMethod method,
            SchwarzFactorEvaluator&& evaluator) {
    init(s1, s2, ln_prec_, method, std::forward<SchwarzFactorEvaluator>(evaluator));
  }
  void reset() {
    primpairs.clear();
    AB[0] = AB[1] = AB[2] = 0;
    ln_prec = std::numeric_limits<real_t>::lowest();
    screening_method_ = ScreeningMethod::Invalid;
  }
  void resize(std::size_t max_nprim) {
    if (max_nprim * max_nprim > primpairs.size()) primpairs.resize(max_nprim * max_nprim);
  }

  /// Original and Conservative screening (shell.h:1137-1256)
  template <typename Real>
  void init(const Shell& s1, const Shell& s2, Real ln_prec_, ScreeningMethod method = ScreeningMethod::Original) {
    if (method != ScreeningMethod::Original && method != ScreeningMethod::Conservative)
      throw std::invalid_argument("ShellPair::init: Original or Conservative screening (Schwarz needs an evaluator)");
    build(s1, s2, (real_t)ln_prec_, method, nullptr);
  }
  /// Schwarz and SchwarzInf screening of primitives (shell.h:1258-1328): the
  /// evaluator returns the Schwarz factor of primitive pair (p1, p2)
  template <typename Real, typename SchwarzFactorEvaluator>
  void init(const Shell& s1, const Shell& s2, Real ln_prec_, ScreeningMethod method,
            SchwarzFactorEvaluator&& evaluator) {
    if (method != ScreeningMethod::Schwarz && method != ScreeningMethod::SchwarzInf)
      throw std::invalid_argument("ShellPair::init: an evaluator goes with Schwarz or SchwarzInf screening");
    std::vector<real_t> f(s1.nprim() * s2.nprim());
    for (size_t p1 = 0; p1 < s1.nprim(); ++p1)
      for (size_t p2 = 0; p2 < s2.nprim(); ++p2) f[p1 * s2.nprim() + p2] = evaluator(s1, p1, s2, p2);
    build(s1, s2, (real_t)ln_prec_, method, f.data());
  }
  /// Engine::compute2's own pair data when the caller gives none, or when
  /// the given pair was screened at a coarser precision (engine.impl.h:
  /// 1259-1276): the Original formula, screened by the Original and
  /// Conservative rules only
  void init_for_engine(const Shell& s1, const Shell& s2, real_t ln_prec_, ScreeningMethod method) {
    build(s1, s2, ln_prec_, method, nullptr);
  }

 private:
  static real_t fact(int n) {
    real_t r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
  }
  void build(const Shell& s1, const Shell& s2, real_t lnp, ScreeningMethod method, const real_t* schwarz) {
    primpairs.clear();
    const auto& A = s1.O;
    const auto& B = s2.O;
    real_t AB2 = 0;
    for (int i = 0; i < 3; ++i) {
      AB[i] = A[i] - B[i];
      AB2 += AB[i] * AB[i];
    }
    int l1 = 0, l2 = 0;
    for (auto& c : s1.contr) l1 = std::max(l1, c.l);
    for (auto& c : s2.contr) l2 = std::max(l2, c.l);
    const size_t n1 = s1.alpha.size(), n2 = s2.alpha.size();
    for (size_t p1 = 0; p1 < n1; ++p1)
      for (size_t p2 = 0; p2 < n2; ++p2) {
        const real_t a1 = s1.alpha[p1], a2 = s2.alpha[p2];
        const real_t oog = 1 / (a1 + a2);
        const real_t mrab2 = -a1 * a2 * oog * AB2;
        real_t ln_scr;
        if (schwarz) {
          ln_scr = std::log((real_t)(n1 * n2) * schwarz[p1 * n2 + p2]) + s1.max_ln_coeff[p1] + s2.max_ln_coeff[p2];
          if (ln_scr < lnp) continue;
        } else {
          ln_scr = mrab2 + s1.max_ln_coeff[p1] + s2.max_ln_coeff[p2];
          if (method == ScreeningMethod::Original && ln_scr < lnp) continue;
        }
        PrimPairData d;
        for (int i = 0; i < 3; ++i) d.P[i] = AB2 == 0 ? A[i] : (a1 * A[i] + a2 * B[i]) * oog;
        d.nonsph_screen_fac = 0;
        if (!schwarz && method == ScreeningMethod::Conservative) {
          real_t pa = 0, pb = 0;
          for (int i = 0; i < 3; ++i) {
            pa = std::max(pa, std::abs(d.P[i] - A[i]));
            pb = std::max(pb, std::abs(d.P[i] - B[i]));
          }
          d.nonsph_screen_fac =
              std::max(std::pow(pa, l1) * std::pow(pb, l2), fact(l1) * fact(l2) * std::pow(oog, l1 + l2));
          // ln(sqrt(2) pi^(5/4)) + ln(1/gamma) + ln(max(nonsph, 1)) + ln(n1 n2)
          ln_scr += 1.777485947591722872387900 + std::log(oog) +
                    std::log(std::max(d.nonsph_screen_fac, (real_t)1)) + std::log((real_t)(n1 * n2));
          if (ln_scr < lnp) continue;
        }
        d.ln_scr = ln_scr;
        d.p1 = (int)p1;
        d.p2 = (int)p2;
        d.K = 5.9149671727956128778 * std::exp(mrab2) * oog;
        d.one_over_gamma = oog;
        primpairs.push_back(d);
      }
    ln_prec = lnp;
    screening_method_ = method;
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
      if (s->ncontr() != 1) throw std::invalid_argument("generally-contracted shells are not yet supported");
      if (s->contr[0].l > lmax_) throw lmax_exceeded("coulomb", lmax_, s->contr[0].l);
    }
    const int method = method_for_plan();  // Original / Conservative without precomputed pairs
    if (deriv_order_ == 0) return compute_resident(sh, nullptr, nullptr);
    // first derivatives: a one-quartet derivative plan
    if (uniform_ || scale_ != 1.0)
      throw std::invalid_argument("libint2_amd: deriv_order 1 with prescaling or uniform normalization is not provided");
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
    l2g_deriv1_plan dp = nullptr;
    detail::check(l2g_eri_deriv1_plan_create(b, 1, q, precision_, method, &dp));
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

  // one quartet on the resident evaluator: one copy in, one launch per
  // operator pass, one copy out.  ln_scr / nonsph: precomputed pair data
  // (bra n1 n2 then ket n3 n4 values, -inf for dropped primitive pairs), or
  // null for pairs built with the Engine's own screening
  const target_ptr_vec& compute_resident(const Shell* const sh[4], const double* ln_scr, const double* nonsph) {
    int l[4], pure[4], np[4], maxnp = 0, maxl = 0;
    double o[12];
    size_t n = 1;
    a_.clear();
    c_.clear();
    for (int k = 0; k < 4; ++k) {
      l[k] = sh[k]->contr[0].l;
      pure[k] = sh[k]->contr[0].pure ? 1 : 0;
      np[k] = (int)sh[k]->nprim();
      maxnp = std::max(maxnp, np[k]);
      maxl = std::max(maxl, l[k]);
      a_.insert(a_.end(), sh[k]->alpha.begin(), sh[k]->alpha.end());
      c_.insert(c_.end(), sh[k]->contr[0].coeff.begin(), sh[k]->contr[0].coeff.end());
      for (int x = 0; x < 3; ++x) o[3 * k + x] = sh[k]->O[x];
      n *= sh[k]->size();
    }
    auto& rt = detail::Runtime::get();
    detail::check(l2g_set_solid_harmonics_ordering(rt.ctx, rt.ordering));
    if (!qe_.h || maxnp > qe_.nprim || maxl > qe_.l) {
      qe_.reset();
      qe_.nprim = std::max<int>((int)max_nprim_, maxnp);
      qe_.l = std::max(lmax_, maxl);
      detail::check(l2g_quartet_engine_create(rt.ctx, qe_.nprim, qe_.l, &qe_.h));
    }
    buffer_.resize(std::max<size_t>(n, 1));
    int8_t nz = 0;
    detail::check(l2g_quartet_compute(qe_.h, l, pure, np, a_.data(), c_.data(), o, ln_scr, nonsph, precision_,
                                      (int)screening_, scale_, uniform_ ? 1 : 0, plan_operator(), omega_,
                                      buffer_.data(), &nz));
    targets_[0] = nz ? buffer_.data() : nullptr;
    return targets_;
  }


 public:
  /// compute2<oper, braket, deriv_order> (engine.h:787-791,
  /// engine.impl.h:1151-2113) for oper coulomb, erf_coulomb or erfc_coulomb,
  /// every 2-body braket (3- and 2-centre with Shell::unit() in the missing
  /// positions, as the caller passes it) and deriv_order 0 or 1.  With
  /// precomputed shell pairs (both or neither, for (bra1, bra2) and
  /// (ket1, ket2) in the caller's order) the primitive pairs and their
  /// screening factors are the ShellPairs', and every screening method is
  /// available; a pair screened at a coarser precision than the Engine's is
  /// rebuilt, as libint does (engine.impl.h:1259-1276).
  template <Operator op, BraKet bk, size_t der>
  const target_ptr_vec& compute2(const Shell& bra1, const Shell& bra2, const Shell& ket1, const Shell& ket2,
                                 const ShellPair* spbra = nullptr, const ShellPair* spket = nullptr) {
    static_assert((op == Operator::coulomb || op == Operator::erf_coulomb || op == Operator::erfc_coulomb) &&
                      bk != BraKet::x_x && bk != BraKet::invalid && der <= 1,
                  "libint2_amd provides compute2 for coulomb, erf_coulomb and erfc_coulomb, 2-body brakets, "
                  "deriv_order 0 and 1");
    if (default_) throw using_default_initialized();
    if (op != oper_ || bk != braket_ || (int)der != deriv_order_)
      throw std::invalid_argument("Engine::compute2 -- operator, braket or deriv_order differs from the Engine's");
    if ((spbra == nullptr) != (spket == nullptr))
      throw std::invalid_argument("Engine::compute2 -- expects zero or two ShellPair objects");
    if (!spbra) return compute_quartet(bra1, bra2, ket1, ket2);
    const Shell* sh[4] = {&bra1, &bra2, &ket1, &ket2};
    for (auto* x : sh) {
      if (x->ncontr() != 1) throw std::invalid_argument("generally-contracted shells are not yet supported");
      if (x->contr[0].l > lmax_) throw lmax_exceeded("coulomb", lmax_, x->contr[0].l);
    }
    if (der == 1)
      throw std::invalid_argument("libint2_amd: compute2 with precomputed shell pairs is provided for deriv_order 0");
    const ShellPair::real_t lnp =
        precision_ > 0 ? std::log(precision_) : std::numeric_limits<ShellPair::real_t>::lowest();
    const ShellPair* sp[2] = {spbra, spket};
    ShellPair own[2];
    ln_scr_.clear();
    nonsph_.clear();
    for (int h = 0; h < 2; ++h) {
      const Shell& s1 = *sh[2 * h];
      const Shell& s2 = *sh[2 * h + 1];
      const ShellPair* p = sp[h];
      if (p->ln_prec > lnp) {  // screened coarser than this Engine's precision
        own[h].init_for_engine(s1, s2, lnp, screening_);
        p = &own[h];
      } else if (p->screening_method_ != screening_) {
        throw std::invalid_argument(
            "Engine::compute2: received ShellPair initialized for an incompatible screening method");
      }
      const size_t n2 = s2.nprim(), base = ln_scr_.size();
      ln_scr_.resize(base + s1.nprim() * n2, -std::numeric_limits<double>::infinity());
      nonsph_.resize(ln_scr_.size(), 0.0);
      for (const auto& pp : p->primpairs) {
        if (pp.p1 < 0 || pp.p2 < 0 || (size_t)pp.p1 >= s1.nprim() || (size_t)pp.p2 >= n2)
          throw std::invalid_argument("Engine::compute2: ShellPair does not match its shells");
        ln_scr_[base + (size_t)pp.p1 * n2 + pp.p2] = pp.ln_scr;
        nonsph_[base + (size_t)pp.p1 * n2 + pp.p2] = pp.nonsph_screen_fac;
      }
    }
    return compute_resident(sh, ln_scr_.data(), nonsph_.data());
  }

  typedef const target_ptr_vec& (Engine::*compute2_ptr_type)(const Shell& bra1, const Shell& bra2,
                                                            const Shell& ket1, const Shell& ket2,
                                                            const ShellPair* spbra, const ShellPair* spket);
  /// the dispatch table of compute2 instantiations, indexed as
  /// ((oper - first_2body_oper) * nbrakets_2body + (braket - first_2body_braket))
  /// * nderivorders_2body + deriv_order (engine.h:1021-1026,
  /// engine.impl.h:150-171, 717-757); null where this build has no
  /// instantiation
  const std::vector<compute2_ptr_type>& compute2_ptrs() const {
    static const std::vector<compute2_ptr_type> tab = [] {
      std::vector<compute2_ptr_type> t(nopers_2body * nbrakets_2body * nderivorders_2body, nullptr);
      auto at = [&](Operator o, BraKet b, size_t d) -> compute2_ptr_type& {
        return t[((size_t)((int)o - (int)Operator::first_2body_oper) * nbrakets_2body +
                  (size_t)((int)b - (int)BraKet::first_2body_braket)) *
                     nderivorders_2body +
                 d];
      };
#define L2G_FACADE_PTRS(O)                                                   \
  at(O, BraKet::xx_xx, 0) = &Engine::compute2<O, BraKet::xx_xx, 0>;          \
  at(O, BraKet::xs_xx, 0) = &Engine::compute2<O, BraKet::xs_xx, 0>;          \
  at(O, BraKet::xx_xs, 0) = &Engine::compute2<O, BraKet::xx_xs, 0>;          \
  at(O, BraKet::xs_xs, 0) = &Engine::compute2<O, BraKet::xs_xs, 0>;
      L2G_FACADE_PTRS(Operator::coulomb)
      L2G_FACADE_PTRS(Operator::erf_coulomb)
      L2G_FACADE_PTRS(Operator::erfc_coulomb)
#undef L2G_FACADE_PTRS
      at(Operator::coulomb, BraKet::xx_xx, 1) = &Engine::compute2<Operator::coulomb, BraKet::xx_xx, 1>;
      return t;
    }();
    return tab;
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
    // prescale_by multiplies the one-body integrals too (engine.impl.h:1050);
    // uniform normalization rescales the Cartesian functions of each shell
    // (engine.impl.h:552-558, cartesian.h:115-133)
    for (size_t i = 0; i < n1; ++i) {
      const double si = uniform_ && !s1.contr[0].pure ? uniform_cart_scale(s1.contr[0].l, (int)i) : 1.0;
      for (size_t j = 0; j < n2; ++j) {
        const double sj = uniform_ && !s2.contr[0].pure ? uniform_cart_scale(s2.contr[0].l, (int)j) : 1.0;
        buffer_[i * n2 + j] = scale_ * si * sj * M[i * n + n1 + j];
      }
    }
    targets_[0] = buffer_.data();
    return targets_;
  }
  // sqrt((2l-1)!! / ((2i-1)!! (2j-1)!! (2k-1)!!)) of Cartesian component c of
  // level l, components in libint's standard order (cartesian.h:96-109)
  static double uniform_cart_scale(int l, int c) {
    int x = l, rem = c;
    while (rem > l - x) {
      rem -= l - x + 1;
      --x;
    }
    const int y = l - x - rem, z = l - x - y;
    auto dfact = [](int k) {
      double r = 1;
      for (int i = 2 * k - 1; i > 1; i -= 2) r *= i;
      return r;
    };
    return std::sqrt(dfact(l) / (dfact(x) * dfact(y) * dfact(z)));
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
  std::vector<double> a_, c_, ln_scr_, nonsph_;  // staging of one call
  // the resident single-quartet evaluator, made on first use; each copy of
  // an Engine (one per thread, engine.h:560-579) makes its own
  struct Resident {
    l2g_quartet_engine h = nullptr;
    int nprim = 0, l = 0;
    Resident() = default;
    Resident(const Resident&) {}
    Resident& operator=(const Resident&) { return *this; }
    Resident(Resident&& o) noexcept : h(o.h), nprim(o.nprim), l(o.l) { o.h = nullptr; }
    Resident& operator=(Resident&& o) noexcept {
      std::swap(h, o.h);
      std::swap(nprim, o.nprim);
      std::swap(l, o.l);
      return *this;
    }
    ~Resident() { reset(); }
    void reset() {
      if (h) l2g_quartet_engine_destroy(h);
      h = nullptr;
    }
  } qe_;
};

}  // namespace libint2
