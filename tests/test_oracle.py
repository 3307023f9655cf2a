"""CPU tests that pin the oracle (the checker of the GPU path).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails at least one of
them: golden vectors from the reference's own tests, an independent
brute-force algorithm, closed forms and symmetry invariants.
"""
import itertools
import json
import os

import numpy as np
import pytest

import libint_amd as la
import oracle
from tests.helpers import compare, random_shells

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = np.finfo(float).eps


class _Arr:
    def __init__(self, shells):
        self.l = [s[0] for s in shells]
        self.pure = [int(s[3]) for s in shells]
        self.nprim = [len(s[1]) for s in shells]
        al, co, org = [], [], []
        for (l, prims, O, pure) in shells:
            a = np.array([p[0] for p in prims], float)
            c, _ = oracle.shell_renorm(l, a, np.array([p[1] for p in prims], float))
            al += list(a)
            co += list(c)
            org += list(O)
        self.alpha, self.coeff, self.origin = al, co, org


def rand_shells(rng, ls, kmax=3, box=(0.7, 1.3), expo=(0.1, 3.0), pure_from=2):
    out = []
    for l in ls:
        K = int(rng.integers(1, kmax + 1))
        prims = list(zip(rng.uniform(*expo, K), rng.uniform(0.1, 1.0, K)))
        out.append((l, prims, tuple(rng.uniform(*box, 3)), l >= pure_from))
    return _Arr(out)


# ---------------------------------------------------------------------------
# Boys function (reference: tests/unit/test-core-ints.cc:28-90)
# ---------------------------------------------------------------------------

def _golden():
    g = json.load(open(os.path.join(HERE, "golden", "boys_golden.json")))
    return np.array(g["T"]), np.array(g["F"])


def test_boys_cheb7_vs_mpmath():
    """abs error <= eps and rel error <= 125 eps, m <= 40, the reference's
    tolerance for FmEval_Chebyshev7 (test-core-ints.cc:79-84)."""
    T, F = _golden()
    v = oracle.boys_cheb7(T, 40)
    d = np.abs(v - F)
    assert (d <= EPS).all()
    assert (d / F <= 125 * EPS).all()


def test_boys_reference_vs_mpmath():
    T, F = _golden()
    v = oracle.boys_reference(T, 40)
    assert (np.abs(v - F) / F <= 4 * EPS).all()


def test_boys_closed_forms():
    """F_0(T) = sqrt(pi/T) erf(sqrt(T)) / 2 and F_m(0) = 1/(2m+1)."""
    from math import erf, pi, sqrt
    T = np.array([1e-3, 0.37, 2.5, 11.0, 60.0, 116.5, 130.0])
    v = oracle.boys_cheb7(T, 0)[:, 0]
    ref = np.array([0.5 * sqrt(pi / t) * erf(sqrt(t)) for t in T])
    assert np.allclose(v, ref, rtol=20 * EPS, atol=0)
    z = oracle.boys_cheb7(np.array([0.0]), 40)[0]
    assert np.allclose(z, 1.0 / (2 * np.arange(41) + 1), rtol=4 * EPS, atol=0)


# ---------------------------------------------------------------------------
# shells and solid harmonics (tests/unit/test-core.cc)
# ---------------------------------------------------------------------------

def test_shell_renorm_known_answer():
    c, mlc = oracle.shell_renorm(2, [1.0], [1.0])
    assert c[0] == pytest.approx(1.64592278064949, rel=1e-13)
    assert mlc[0] == pytest.approx(np.log(c[0]))


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4])
def test_shell_unit_self_overlap(l):
    """The contraction is rescaled to unit self-overlap of x^l
    (shell.h:976-994): sum_pq c_p c_q (2l-1)!! pi^{3/2} / (2^l g^{l+3/2}) = 1."""
    a = np.array([0.2, 1.3, 7.0])
    c, _ = oracle.shell_renorm(l, a, [0.3, 0.5, 0.9])
    dfl = np.prod(np.arange(2 * l - 1, 0, -2)) if l > 0 else 1.0
    s = sum(c[p] * c[q] * dfl * np.pi ** 1.5 / (2 ** l * (a[p] + a[q]) ** (l + 1.5))
            for p in range(3) for q in range(3))
    assert s == pytest.approx(1.0, abs=1e-14)


def test_solid_harmonics_known_answers():
    # Catch's default Approx margin, 100 float epsilons (test-core.cc:26-37);
    # evaluated here in double, where l = 25 loses ~8 digits to factorials
    assert oracle.solidharm_coeff(12, 0, 0, 0, 12) == pytest.approx(1.0, abs=1e-12)
    assert oracle.solidharm_coeff(25, 0, 0, 0, 25) == pytest.approx(1.0, rel=1.2e-5)
    # pure p: m=+1 <-> x, m=-1 <-> y, m=0 <-> z (tests/unit/test-shell-order.cc:52-72)
    assert oracle.solidharm_coeff(1, 1, 1, 0, 0) == pytest.approx(1.0)
    assert oracle.solidharm_coeff(1, -1, 0, 1, 0) == pytest.approx(1.0)
    assert oracle.solidharm_coeff(1, 0, 0, 0, 1) == pytest.approx(1.0)
    assert oracle.solidharm_coeff(1, 1, 0, 1, 0) == 0.0


def _cart(l):
    return [(i, j, l - i - j) for i in range(l, -1, -1) for j in range(l - i, -1, -1)]


@pytest.mark.parametrize("l", [2, 3, 4])
def test_solid_harmonics_orthonormal(l):
    """With libint's Cartesian normalization (x^l normalised, so
    <x^a y^b z^c | x^a' ...> = (a+a'-1)!!(b+b'-1)!!(c+c'-1)!!/(2l-1)!! for
    even sums), the real solid harmonics are orthonormal: C S C^T = 1."""
    def df(n):  # (n)!! for odd n, 1 for n <= 0
        return float(np.prod(np.arange(n, 0, -2))) if n > 0 else 1.0
    cc = _cart(l)
    S = np.zeros((len(cc), len(cc)))
    for i, a in enumerate(cc):
        for j, b in enumerate(cc):
            s = [a[k] + b[k] for k in range(3)]
            if any(x % 2 for x in s):
                continue
            S[i, j] = np.prod([df(x - 1) for x in s]) / df(2 * l - 1)
    C = np.array([[oracle.solidharm_coeff(l, m, *c) for c in cc] for m in range(-l, l + 1)])
    assert np.allclose(C @ S @ C.T, np.eye(2 * l + 1), atol=1e-13)


# ---------------------------------------------------------------------------
# ERIs: goldens, brute force, symmetry
# ---------------------------------------------------------------------------

def test_python_goldens():
    """python/tests/test_libint2.py:36-45 with s = Shell(0,[(1,10)]),
    p = Shell(1,[(1,10)]) at the origin, pure=True (libint2.cc:155-156)."""
    b = _Arr([(0, [(1.0, 10.0)], (0, 0, 0), True), (1, [(1.0, 10.0)], (0, 0, 0), True)])
    v, _, _ = oracle.eri_quartets(b, [[1, 1, 0, 0]])
    assert np.linalg.norm(v) == pytest.approx(1.62867503968, abs=5e-11)
    b4 = _Arr([(0, [(1.0, 10.0)], (0, 0, 0), True), (1, [(1.0, 10.0)], (0, 0, 0), True)] * 2)
    q = list(itertools.product(range(4), repeat=4))
    v, _, _ = oracle.eri_quartets(b4, q)
    assert np.linalg.norm(v) == pytest.approx(14.7036075402, abs=5e-10)
    vb = oracle.eri_brute(b4, q)
    assert np.linalg.norm(vb) == pytest.approx(14.7036075402, abs=5e-10)


def test_ssss_closed_form():
    """(ss|ss) over normalised primitives has the closed form
    2 pi^{5/2} / (pq sqrt(p+q)) exp(-mu_ab AB^2 - mu_cd CD^2) F_0(T) N^4."""
    from math import erf, exp, pi, sqrt
    al = [0.7, 1.9, 0.4, 3.3]
    O = [(0, 0, 0), (0.5, 0.1, -0.3), (1.0, -0.4, 0.2), (-0.2, 0.9, 0.6)]
    b = _Arr([(0, [(a, 1.0)], o, False) for a, o in zip(al, O)])
    v, _, _ = oracle.eri_quartets(b, [[0, 1, 2, 3]], precision=0.0)
    A, B, C, D = (np.array(o) for o in O)
    p, q = al[0] + al[1], al[2] + al[3]
    P = (al[0] * A + al[1] * B) / p
    Q = (al[2] * C + al[3] * D) / q
    T = p * q / (p + q) * np.dot(P - Q, P - Q)
    F0 = 0.5 * sqrt(pi / T) * erf(sqrt(T))
    N = np.prod([(2 * a / pi) ** 0.75 for a in al])
    ref = N * 2 * pi ** 2.5 / (p * q * sqrt(p + q)) * exp(
        -al[0] * al[1] / p * np.dot(A - B, A - B) - al[2] * al[3] / q * np.dot(C - D, C - D)) * F0
    assert v[0] == pytest.approx(ref, rel=1e-13)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_vs_brute_force(seed):
    """Every class up to (dd|dd) and mixed orders, contracted K<=2, against the
    McMurchie-Davidson brute force; fail iff abs > 5e-14 and rel > 1e-9
    (tests/eri/test.cc:77-88)."""
    rng = np.random.default_rng(seed)
    b = rand_shells(rng, [0, 1, 2, 0, 1, 2], kmax=2)
    q = np.array(list(itertools.product(range(6), repeat=4)), np.int32)
    q = q[rng.choice(len(q), 150, replace=False)]
    v, _, _ = oracle.eri_quartets(b, q, precision=0.0)
    vb = oracle.eri_brute(b, q)
    nfail, mabs, _ = compare(v, vb, 5e-14, 1e-9)
    assert nfail == 0, (nfail, mabs)


def test_oracle_vs_brute_force_f():
    """f shells (K=1) against the brute force."""
    rng = np.random.default_rng(9)
    b = rand_shells(rng, [3, 1, 2, 0], kmax=1)
    q = [[0, 1, 0, 1], [0, 3, 2, 1], [2, 0, 3, 3], [1, 1, 0, 2]]
    v, _, _ = oracle.eri_quartets(b, q, precision=0.0)
    vb = oracle.eri_brute(b, q)
    nfail, _, _ = compare(v, vb, 5e-14, 1e-9)
    assert nfail == 0


def test_oracle_ld_vs_brute_force():
    """The long-double restatement (the arbiter of the GPU parity tests)
    against the McMurchie-Davidson brute force, classes to (ff|ff) with K = 2,
    at 1e-15 absolute: both carry ~19 significant digits."""
    rng = np.random.default_rng(11)
    b = rand_shells(rng, [3, 3, 2, 1], kmax=2)
    q = [[0, 1, 0, 1], [1, 0, 3, 2], [0, 0, 1, 1], [2, 3, 1, 0], [3, 2, 0, 0]]
    vb = oracle.eri_brute(b, q)
    vl, _, _ = oracle.eri_quartets_ld(b, q, precision=0.0)
    assert np.abs(vl - vb).max() <= 1e-15


def test_oracle_ld_resolves_hrr_cancellation():
    """A distant pair of a diffuse and a tight d shell built on the diffuse
    one (the caller's order for equal l, as the reference builds it): the
    HRR cancels, the double restatement is off the brute force by more than
    the integrals themselves, the long-double one is not."""
    b = _Arr([(2, [(0.06, 1.0)], (0, 0, 0), True), (2, [(3000.0, 1.0)], (4.0, 1.0, -2.0), True),
              (2, [(0.08, 1.0)], (0.5, 0, 0), True), (2, [(900.0, 1.0)], (3.0, 2.5, -1.0), True)])
    q = [[0, 1, 2, 3]]
    vb = oracle.eri_brute(b, q)
    vd, _, _ = oracle.eri_quartets(b, q, precision=0.0)
    vl, _, _ = oracle.eri_quartets_ld(b, q, precision=0.0)
    assert np.abs(vd - vb).max() > np.abs(vb).max()  # the double restatement's own error
    assert np.abs(vl - vb).max() <= 1e-15


def test_oracle_quad_vs_brute_force_ill_conditioned():
    """The parity arbiter: the restatement in quad precision against the
    brute force on a distant (ff|ff) pair whose contractions mix diffuse and
    tight primitives on both shells.  There the reference's FP64 algorithm is
    off by more than 1e-6 (5e-4 of the largest integral) and long double by
    more than 1e-9 in either build orientation; quad precision matches the
    brute force to 1e-17."""
    b = _Arr([(3, [(0.04, 0.6), (3000.0, 0.5)], (0, 0, 0), True),
              (3, [(0.05, 0.7), (2500.0, 0.4)], (9.0, 0.7, -0.4), True)])
    q = [[0, 1, 0, 1]]
    vb = oracle.eri_brute(b, q)
    vd, _, _ = oracle.eri_quartets(b, q, precision=0.0)
    vl, _, _ = oracle.eri_quartets_ld(b, q, precision=0.0, tight_centre=True)
    vq, _, _ = oracle.eri_quartets_quad(b, q, precision=0.0)
    assert np.abs(vd - vb).max() > 1e-6
    assert np.abs(vl - vb).max() > 1e-9
    assert np.abs(vq - vb).max() <= 1e-17


def test_oracle_quad_matches_double():
    """On well-conditioned quartets quad and double agree to 1e-13, with the
    same layout and screened-out flags."""
    rng = np.random.default_rng(13)
    b = rand_shells(rng, [0, 1, 2, 3, 2, 1], kmax=2)
    q = np.array(list(itertools.product(range(6), repeat=4)), np.int32)
    q = q[rng.choice(len(q), 60, replace=False)]
    vd, od, nd = oracle.eri_quartets(b, q)
    vq, oq, nq = oracle.eri_quartets_quad(b, q)
    assert np.array_equal(od, oq) and np.array_equal(nd, nq)
    assert np.abs(vd - vq).max() <= 1e-13


def test_oracle_ld_matches_double():
    """On well-conditioned quartets (random shells near each other, K <= 3)
    the two precisions agree to 1e-13."""
    rng = np.random.default_rng(12)
    b = rand_shells(rng, [0, 1, 2, 3, 2, 1], kmax=3)
    q = np.array(list(itertools.product(range(6), repeat=4)), np.int32)
    q = q[rng.choice(len(q), 200, replace=False)]
    vd, od, nd = oracle.eri_quartets(b, q)
    vl, ol, nl = oracle.eri_quartets_ld(b, q)
    assert np.array_equal(od, ol) and np.array_equal(nd, nl)
    assert np.abs(vd - vl).max() <= 1e-13


def test_c_api_case_vs_brute_force():
    """The primitive (pp|pp) case of tests/unit/test-c-api.cc:40-96."""
    b = _Arr([(1, [(1.1, 1.0)], (0, 1, 2), False), (1, [(2.3, 1.0)], (1, 2, 0), False),
              (1, [(3.4, 1.0)], (2, 0, 1), False), (1, [(4.8, 1.0)], (0, 1, 2), False)])
    v, _, _ = oracle.eri_quartets(b, [[0, 1, 2, 3]], precision=0.0)
    vb = oracle.eri_brute(b, [[0, 1, 2, 3]])
    nfail, _, _ = compare(v, vb, 5e-14, 1e-9)
    assert nfail == 0


def test_permutational_symmetry():
    """All 8 permutations of a (dp|sd) quartet agree to 1e-13
    (tests/unit/test-permute.cc:90-178, :116)."""
    rng = np.random.default_rng(3)
    b = rand_shells(rng, [2, 1, 0, 2], kmax=3)
    sizes = [5, 3, 1, 5]
    base, _, _ = oracle.eri_quartets(b, [[0, 1, 2, 3]])
    ref = base.reshape(sizes)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2), (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0)]:
        v, _, _ = oracle.eri_quartets(b, [list(perm)])
        t = v.reshape([sizes[p] for p in perm])
        # axis k of t is shell perm[k]; bring back to (0,1,2,3)
        back = np.transpose(t, np.argsort(perm))
        assert np.abs(back - ref).max() < 1e-13


def test_braket_swap_uniform_normalization():
    """(12|34) vs (34|12) with uniform Cartesian normalization, abs < 1e-14
    (tests/unit/test-2body.cc:126-189)."""
    rng = np.random.default_rng(4)
    b = rand_shells(rng, [1, 1, 1, 1], kmax=2, pure_from=9)
    v1, _, _ = oracle.eri_quartets(b, [[0, 1, 2, 3]], uniform_cart=True)
    v2, _, _ = oracle.eri_quartets(b, [[2, 3, 0, 1]], uniform_cart=True)
    a = v1.reshape(9, 9)
    c = v2.reshape(9, 9).T
    assert np.abs(a - c).max() < 1e-14


def test_uniform_normalization_self_overlap_pattern():
    """Standard normalization gives x^a y^b z^c diagonal ratios
    (2a-1)!!(2b-1)!!(2c-1)!!/(2l-1)!! (test-core.cc:92-108); uniform scaling
    multiplies each Cartesian by its inverse square root.  Check through
    (d s|d s)-type diagonal elements with a tight ket."""
    b = _Arr([(2, [(1.0, 1.0)], (0, 0, 0), False), (0, [(1.0, 1.0)], (0, 0, 0), False)])
    std, _, _ = oracle.eri_quartets(b, [[0, 1, 0, 1]], precision=0.0)
    uni, _, _ = oracle.eri_quartets(b, [[0, 1, 0, 1]], precision=0.0, uniform_cart=True)
    ratio = np.diag(uni.reshape(6, 6)) / np.diag(std.reshape(6, 6))
    ref = np.array([1, 3, 3, 1, 3, 1], float)  # 3!!/(1!!1!!) for xy-type, 1 for axial
    assert np.allclose(ratio, ref, rtol=1e-13)


def test_screening_null_result():
    """All primitives screened -> nonzero flag 0 and zeros
    (engine.impl.h:1781-1784)."""
    b = _Arr([(0, [(80.0, 1.0)], (0, 0, 0), False), (0, [(80.0, 1.0)], (6.0, 0, 0), False)])
    v, _, nz = oracle.eri_quartets(b, [[0, 1, 0, 1], [0, 0, 0, 0]], precision=1e-10)
    assert nz[0] == 0 and v[0] == 0.0
    assert nz[1] == 1 and v[1] > 0


def test_oracle_three_centre_golden():
    """3-centre (s s|s) through Shell.unit() (exponent 0, coefficient 1) as
    the missing centre: the reference's python golden
    ||(s s|s)|| = 3.6563211198 for s = Shell(0, [(1, 10)])
    (python/tests/test_libint2.py:39-42)."""
    s = la.Shell(0, [(1.0, 10.0)])
    arrays = la.BasisArrays.from_shells([s, la.Shell.unit()])
    v, _, _ = oracle.eri_quartets(arrays, [[0, 1, 0, 0]], precision=0.0)
    assert abs(np.linalg.norm(v) - 3.6563211198) < 1e-9
    # and the 2-centre (p|p) equals the 4-centre integral with unit partners
    p = la.Shell(1, [(1.3, 0.5), (0.4, 0.6)], (0.2, 0.0, 0.1))
    arrays = la.BasisArrays.from_shells([p, la.Shell.unit(), p])
    v2, _, _ = oracle.eri_quartets(arrays, [[0, 1, 2, 1]], precision=0.0)
    v3, _, _ = oracle.eri_quartets(arrays, [[1, 0, 1, 2]], precision=0.0)
    assert np.allclose(v2, v3.reshape(3, 3).T.reshape(-1), rtol=1e-13, atol=1e-15)


def test_oracle_range_separated_ss_quadrature():
    """erf_coulomb (ss|ss) against an independent quadrature of
    erf(w r)/r = 2/sqrt(pi) int_0^w exp(-t^2 r^2) dt over closed-form
    Gaussian-geminal integrals; erfc_coulomb as Coulomb minus it."""
    from scipy.special import roots_legendre
    rng = np.random.default_rng(51)
    for _ in range(4):
        ex = rng.uniform(0.2, 3.0, 4)
        cen = rng.uniform(-1, 1, (4, 3))
        sh = [la.Shell(0, [(x, 1.0)], tuple(o)) for x, o in zip(ex, cen)]
        arr = la.BasisArrays.from_shells(sh)
        a, b, c, d = ex
        A, B, C, D = cen
        p, q = a + b, c + d
        P, Q = (a * A + b * B) / p, (c * C + d * D) / q
        N = np.prod([(2 * x / np.pi) ** 0.75 for x in ex])
        K = np.exp(-a * b / p * np.sum((A - B) ** 2) - c * d / q * np.sum((C - D) ** 2))
        for w in (0.3, 1.1, 4.0):
            x, wt = roots_legendre(300)
            t, wt = 0.5 * w * (x + 1), 0.5 * w * wt
            den = p * q + t * t * (p + q)
            I = (np.pi ** 2 / (p * q)) ** 1.5 * (p * q / den) ** 1.5 * np.exp(-p * q * t * t / den * np.sum((P - Q) ** 2))
            ref = N * K * 2 / np.sqrt(np.pi) * np.sum(wt * I)
            v, _, _ = oracle.eri_quartets(arr, [[0, 1, 2, 3]], precision=0.0, oper=oracle.ERF_COULOMB, omega=w)
            vc, _, _ = oracle.eri_quartets(arr, [[0, 1, 2, 3]], precision=0.0)
            ve, _, _ = oracle.eri_quartets(arr, [[0, 1, 2, 3]], precision=0.0, oper=oracle.ERFC_COULOMB, omega=w)
            assert abs(v[0] - ref) < 1e-14 + 1e-12 * abs(ref)
            assert abs(ve[0] - (vc[0] - ref)) < 1e-14 + 1e-12 * abs(vc[0])


def test_oracle_range_separated_vs_brute_force():
    """erf_coulomb and erfc_coulomb from the restated OS path against the
    McMurchie-Davidson brute force with the attenuated Hermite exponent, for
    every class up to l = 2 plus f, at the brute-force bar."""
    rng = np.random.default_rng(52)
    shells = random_shells(rng, [0, 1, 2, 3, 1, 2], kmax=2)
    arr = la.BasisArrays.from_shells(shells)
    n = len(shells)
    q = np.array([(i, j, k, l) for i in range(n) for j in range(n) for k in range(n) for l in range(n)
                  if (i * 7 + j * 5 + k * 3 + l) % 11 == 0], np.int32)
    for oper in (oracle.ERF_COULOMB, oracle.ERFC_COULOMB):
        for w in (0.4, 2.5):
            o, _, _ = oracle.eri_quartets(arr, q, precision=0.0, oper=oper, omega=w)
            bf = oracle.eri_brute(arr, q, oper=oper, omega=w)
            assert compare(o, bf, 5e-14, 1e-9)[0] == 0, (oper, w)


def test_parity_arbitration_rule():
    """tests.helpers.compare_arbitrated on the ill-conditioned (ff|ff) pair
    above and on well-conditioned quartets: values with half the
    reference's error on the ill-conditioned quartet pass, their misses
    counted and bounded by the reference's; a value off the exact one by
    more than four times the reference's worst error on that quartet fails,
    and so does a value off the exact one beyond the bar on a quartet the
    reference gets right."""
    from tests.helpers import compare_arbitrated, within_reference
    b = _Arr([(3, [(0.04, 0.6), (3000.0, 0.5)], (0, 0, 0), True),
              (3, [(0.05, 0.7), (2500.0, 0.4)], (9.0, 0.7, -0.4), True)])
    q = np.array([[0, 1, 0, 1]], np.int32)
    ref, off, _ = oracle.eri_quartets(b, q, precision=0.0)
    exact, _, _ = oracle.eri_quartets_quad(b, q, precision=0.0)
    scale = np.abs(ref - exact).max()
    assert scale > 1e-6
    half = exact + 0.5 * (ref - exact)
    det = {}
    nfail, narb, oerr = compare_arbitrated(b, q, off, half, ref, 1e-12, 1e-10, detail=det)
    assert nfail == 0 and narb > 0 and oerr == scale
    assert det["beyond_bar"] > 0 and det["beyond_bar_max"] <= 0.5 * det["ref_max"] * (1 + 1e-12)
    assert within_reference(det)
    worse = exact.copy()
    worse[0] += 5 * scale
    nfail, _, _ = compare_arbitrated(b, q, off, worse, ref, 1e-12, 1e-10)
    assert nfail == 1
    # a well-conditioned quartet: the bar itself
    rng = np.random.default_rng(17)
    b2 = rand_shells(rng, [1, 2, 1, 0], kmax=2)
    q2 = np.array([[0, 1, 2, 3]], np.int32)
    v2, off2, _ = oracle.eri_quartets(b2, q2)
    bad = v2.copy()
    k = int(np.argmax(np.abs(v2)))
    bad[k] += 1e-9 + 1e-9 * abs(v2[k])
    nfail, narb, _ = compare_arbitrated(b2, q2, off2, bad, v2, 1e-12, 1e-10)
    assert nfail == 1 and narb == 1
    assert compare_arbitrated(b2, q2, off2, v2, v2, 1e-12, 1e-10)[0] == 0


def test_fock_rows_only_matches_full_build():
    """The row-restricted Fock build (quartets touching the sampled shells
    only) gives the full build's rows of G, J, K for those shells' functions,
    and leaves others partial, so the filter is not a no-op."""
    from libint_amd import molecule
    mol = molecule.make_molecule(2, natoms=3, same_density=True)
    rng = np.random.default_rng(3)
    C = rng.normal(0.0, 0.1, (mol.nbf, mol.nbf // 5))
    D = C @ C.T
    G, J, K, nq = oracle.fock_2e(mol.basis, D, 1e-12, nthreads=4)
    shells = [0, 7, int(mol.basis.nshell) - 1]
    rows = oracle.shell_rows(mol.basis, shells)
    with oracle.fock_rows_only(mol.basis, shells):
        Gr, Jr, Kr, nqr = oracle.fock_2e(mol.basis, D, 1e-12, nthreads=4)
    assert 0 < nqr < nq
    for X, Y in ((G, Gr), (J, Jr), (K, Kr)):
        assert np.allclose(X[rows], Y[rows], rtol=0, atol=1e-13)
        assert np.allclose(X[:, rows], Y[:, rows], rtol=0, atol=1e-13)
    others = np.setdiff1d(np.arange(mol.nbf), rows)
    assert np.abs(G[np.ix_(others, others)] - Gr[np.ix_(others, others)]).max() > 1e-3
