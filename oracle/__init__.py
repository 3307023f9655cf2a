"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes front-end for ``oracle/lib/liboracle.so`` (built from
``oracle/eri_oracle.cpp`` by ``oracle/Makefile``).  The oracle is a CPU
restatement of libint's ``Engine::compute2<Operator::coulomb, BraKet::xx_xx, 0>``
(include/libint2/engine.impl.h:1151-2113) plus an independent
McMurchie-Davidson brute force and long-double Boys reference.  Only
``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module; the shipped library never does.

Parity status per function (how each is pinned is listed in the tests):
  * boys_cheb7 / boys_reference: pinned against mpmath goldens
    (tests/golden/boys_golden.json), tolerance of
    tests/unit/test-core-ints.cc:79-84.  Bit content of libint's own
    Chebyshev table is UNPINNED (the table is absent from the reference tree).
  * shell_renorm: known answer 1.64592278064949 (tests/unit/test-core.cc:49-55)
    and the unit self-overlap invariant.
  * solidharm_coeff: known answers of tests/unit/test-core.cc:26-37 and the
    p-shell mapping of tests/unit/test-shell-order.cc:52-72.
  * eri_quartets: brute force (eri_brute), the Python goldens of
    python/tests/test_libint2.py:36-45, 8-fold permutational symmetry.
  * eri_brute: closed forms ((ss|ss) analytic), the same Python goldens.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

ORIGINAL, CONSERVATIVE, SCHWARZ, SCHWARZ_INF = 0x0001, 0x0010, 0x0100, 0x1000

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile the oracle library (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_native_build(outdir: str) -> str:
    """Compiles the oracle for the ISA of the host this runs on
    (g++ -O3 -march=native) into outdir and loads that library instead of
    the portable build (x86-64-v3, made where the repository is built and
    loaded wherever it runs).  For timing the CPU baseline on the machine
    that measures the GPU.  Returns the compile flags used."""
    global _lib, _LIBPATH
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, "liboracle_native.so")
    flags = ["-O3", "-march=native", "-fopenmp", "-fPIC", "-std=c++17", "-fext-numeric-literals"]
    subprocess.run(["g++", *flags, "-shared", "-o", path, os.path.join(_HERE, "eri_oracle.cpp"), "-lquadmath"],
                   check=True)
    _lib = None
    _LIBPATH = path
    lib()
    return " ".join(flags)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        L.oracle_set_shg_ordering.argtypes = [ctypes.c_int]
        L.oracle_set_operator.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_set_tight_centre.argtypes = [ctypes.c_int]
        L.oracle_boys_cheb7.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, _f64p]
        L.oracle_boys_reference.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, _f64p]
        L.oracle_boys_table.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_boys_table.restype = ctypes.POINTER(ctypes.c_double)
        L.oracle_shell_renorm.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p, ctypes.c_int]
        L.oracle_solidharm_coeff.argtypes = [ctypes.c_int] * 5
        L.oracle_solidharm_coeff.restype = ctypes.c_double
        basis_args = [ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.oracle_eri_quartets.argtypes = basis_args + [
            ctypes.c_int64, _i32p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
            ctypes.c_void_p, _i64p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_eri_quartets.restype = ctypes.c_int64
        L.oracle_eri_quartets_ld.argtypes = L.oracle_eri_quartets.argtypes
        L.oracle_eri_quartets_ld.restype = ctypes.c_int64
        L.oracle_eri_quartets_q.argtypes = L.oracle_eri_quartets.argtypes
        L.oracle_eri_quartets_q.restype = ctypes.c_int64
        L.oracle_eri_brute.argtypes = basis_args + [ctypes.c_int64, _i32p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_eri_brute.restype = ctypes.c_int64
        L.oracle_fock_2e.argtypes = basis_args + [
            _f64p, ctypes.c_double, _f64p, ctypes.c_int, _f64p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.oracle_fock_2e.restype = ctypes.c_int
        L.oracle_fock_2e_arbitrated.argtypes = basis_args + [
            _f64p, ctypes.c_double, _f64p, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int,
            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.oracle_fock_2e_arbitrated.restype = ctypes.c_int
        L.oracle_set_fock_rows.argtypes = [ctypes.c_int, ctypes.c_void_p]
        L.oracle_fock_pairs.argtypes = basis_args + [ctypes.c_double, _i8p, _f64p]
        L.oracle_fock_pairs.restype = ctypes.c_int
        _lib = L
    return _lib


@dataclass
class OBasis:
    """Flat shell arrays (coefficients already renormalised, bohr)."""
    l: np.ndarray
    pure: np.ndarray
    nprim: np.ndarray
    alpha: np.ndarray
    coeff: np.ndarray
    origin: np.ndarray

    @staticmethod
    def of(b) -> "OBasis":
        return OBasis(np.ascontiguousarray(b.l, np.int32), np.ascontiguousarray(b.pure, np.int32),
                      np.ascontiguousarray(b.nprim, np.int32), np.ascontiguousarray(b.alpha, np.float64),
                      np.ascontiguousarray(b.coeff, np.float64),
                      np.ascontiguousarray(np.asarray(b.origin, np.float64).reshape(-1)))

    def args(self):
        return (len(self.l), self.l, self.pure, self.nprim, self.alpha, self.coeff, self.origin)


def set_shg_ordering(gaussian: bool) -> None:
    lib().oracle_set_shg_ordering(1 if gaussian else 0)


def boys_cheb7(T, mmax: int) -> np.ndarray:
    T = np.ascontiguousarray(np.atleast_1d(T), np.float64)
    out = np.zeros((T.size, mmax + 1))
    lib().oracle_boys_cheb7(T.size, T, mmax, out)
    return out


def boys_reference(T, mmax: int) -> np.ndarray:
    T = np.ascontiguousarray(np.atleast_1d(T), np.float64)
    out = np.zeros((T.size, mmax + 1))
    lib().oracle_boys_reference(T.size, T, mmax, out)
    return out


def boys_table() -> np.ndarray:
    nint, mmax = ctypes.c_int(), ctypes.c_int()
    p = lib().oracle_boys_table(ctypes.byref(nint), ctypes.byref(mmax))
    n = nint.value * (mmax.value + 1) * 8
    return np.ctypeslib.as_array(p, shape=(n,)).reshape(nint.value, mmax.value + 1, 8).copy()


def shell_renorm(l: int, alpha, coeff, enforce_unit: bool = True):
    a = np.ascontiguousarray(alpha, np.float64)
    c = np.array(coeff, np.float64)
    mlc = np.zeros_like(c)
    lib().oracle_shell_renorm(l, a.size, a, c, mlc, 1 if enforce_unit else 0)
    return c, mlc


def solidharm_coeff(l, m, lx, ly, lz) -> float:
    return lib().oracle_solidharm_coeff(l, m, lx, ly, lz)


COULOMB, ERF_COULOMB, ERFC_COULOMB = 0, 1, 2


def eri_quartets(basis, quartets, precision=np.finfo(float).eps, screening=ORIGINAL,
                 uniform_cart=False, nthreads=0, oper=COULOMB, omega=0.0, tight_centre=False):
    """Restated Engine::compute2 over quartets (caller order), for
    Operator::coulomb, erf_coulomb or erfc_coulomb (omega: the attenuation).

    Returns (flat values, offsets[nq], nonzero[nq])."""
    ob = OBasis.of(basis)
    q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
    nq = q.shape[0]
    offs = np.zeros(nq, np.int64)
    L = lib()
    if L.oracle_set_operator(int(oper), float(omega)) != 0:
        raise ValueError("bad operator or omega")
    L.oracle_set_tight_centre(int(bool(tight_centre)))
    try:
        total = L.oracle_eri_quartets(*ob.args(), nq, q.reshape(-1), precision, screening,
                                      int(uniform_cart), None, offs, None, nthreads)
        out = np.zeros(max(total, 1))
        nz = np.zeros(nq, np.int8)
        L.oracle_eri_quartets(*ob.args(), nq, q.reshape(-1), precision, screening, int(uniform_cart),
                              out.ctypes.data, offs, nz.ctypes.data, nthreads)
    finally:
        L.oracle_set_operator(COULOMB, 0.0)
        L.oracle_set_tight_centre(0)
    return out[:total], offs, nz


def eri_quartets_quad(basis, quartets, precision=np.finfo(float).eps, screening=ORIGINAL,
                      uniform_cart=False, nthreads=0, oper=COULOMB, omega=0.0):
    """eri_quartets in quad precision (__float128) throughout, rounded to
    double: keeps ~20 digits where HGP's HRR amplifies rounding by 1e12 and
    more (deeply contracted distant pairs of diffuse and tight primitives),
    where the double and long-double restatements both fail.  Returns (flat
    values, offsets, nonzero)."""
    return _eri_quartets_with("oracle_eri_quartets_q", basis, quartets, precision, screening, uniform_cart,
                              nthreads, oper, omega, False)


def _eri_quartets_with(fn, basis, quartets, precision, screening, uniform_cart, nthreads, oper, omega,
                       tight_centre):
    ob = OBasis.of(basis)
    q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
    nq = q.shape[0]
    offs = np.zeros(nq, np.int64)
    L = lib()
    f = getattr(L, fn)
    if L.oracle_set_operator(int(oper), float(omega)) != 0:
        raise ValueError("bad operator or omega")
    L.oracle_set_tight_centre(int(bool(tight_centre)))
    try:
        total = f(*ob.args(), nq, q.reshape(-1), precision, screening, int(uniform_cart), None, offs, None,
                  nthreads)
        out = np.zeros(max(total, 1))
        nz = np.zeros(nq, np.int8)
        f(*ob.args(), nq, q.reshape(-1), precision, screening, int(uniform_cart), out.ctypes.data, offs,
          nz.ctypes.data, nthreads)
    finally:
        L.oracle_set_operator(COULOMB, 0.0)
        L.oracle_set_tight_centre(0)
    return out[:total], offs, nz


def eri_quartets_ld(basis, quartets, precision=np.finfo(float).eps, screening=ORIGINAL,
                    uniform_cart=False, nthreads=0, oper=COULOMB, omega=0.0, tight_centre=False):
    """eri_quartets evaluated in long double throughout (Boys function by
    its series, pair data from the shells), rounded to double; with
    tight_centre on the tighter shell of each equal-l pair: the arbiter for
    values where the double restatement loses digits to HRR cancellation.
    Returns (flat values, offsets, nonzero)."""
    ob = OBasis.of(basis)
    q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
    nq = q.shape[0]
    offs = np.zeros(nq, np.int64)
    L = lib()
    if L.oracle_set_operator(int(oper), float(omega)) != 0:
        raise ValueError("bad operator or omega")
    L.oracle_set_tight_centre(int(bool(tight_centre)))
    try:
        total = L.oracle_eri_quartets_ld(*ob.args(), nq, q.reshape(-1), precision, screening,
                                         int(uniform_cart), None, offs, None, nthreads)
        out = np.zeros(max(total, 1))
        nz = np.zeros(nq, np.int8)
        L.oracle_eri_quartets_ld(*ob.args(), nq, q.reshape(-1), precision, screening, int(uniform_cart),
                                 out.ctypes.data, offs, nz.ctypes.data, nthreads)
    finally:
        L.oracle_set_operator(COULOMB, 0.0)
        L.oracle_set_tight_centre(0)
    return out[:total], offs, nz


def eri_brute(basis, quartets, nthreads=0, oper=COULOMB, omega=0.0):
    """McMurchie-Davidson in long double (an independent derivation); for
    erf_coulomb the Hermite integrals take the attenuated exponent."""
    ob = OBasis.of(basis)
    q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
    L = lib()
    if L.oracle_set_operator(int(oper), float(omega)) != 0:
        raise ValueError("bad operator or omega")
    try:
        total = L.oracle_eri_brute(*ob.args(), q.shape[0], q.reshape(-1), None, nthreads)
        out = np.zeros(max(total, 1))
        L.oracle_eri_brute(*ob.args(), q.shape[0], q.reshape(-1), out.ctypes.data, nthreads)
    finally:
        L.oracle_set_operator(COULOMB, 0.0)
    return out[:total]


def fock_2e(basis, D, precision=1e-12, screening=SCHWARZ_INF, nthreads=0, tight_centre=False):
    """Restated compute_2body_fock: returns (G, J, K, n_quartets) with
    G = 2J - K (tests/hartree-fock/hartree-fock++.cc:1574-1772).
    tight_centre: every integral (Schwarz factors included) built on the
    tighter shell of an equal-l pair instead of the caller's first one: the
    same quantities with far less HRR cancellation, the accurate reference
    where the reference's own orientation loses digits."""
    ob = OBasis.of(basis)
    D = np.ascontiguousarray(D, np.float64)
    N = D.shape[0]
    G = np.zeros((N, N))
    JK = np.zeros((2, N, N))
    nq = ctypes.c_int64()
    L = lib()
    L.oracle_set_tight_centre(int(bool(tight_centre)))
    try:
        L.oracle_fock_2e(*ob.args(), D.reshape(-1), precision, G.reshape(-1), screening, JK.reshape(-1),
                         nthreads, ctypes.byref(nq))
    finally:
        L.oracle_set_tight_centre(0)
    return G, JK[0], JK[1], nq.value


def set_progress(on: bool) -> None:
    """Long Fock builds print one line per shell row to stderr."""
    lib().oracle_set_progress(int(bool(on)))


def fock_2e_exact(basis, D, precision=1e-12, screening=SCHWARZ_INF, nthreads=0, all_quad=False):
    """fock_2e(tight_centre=True) with the ill-conditioned quartets settled
    in quad precision: every quartet is evaluated in FP64 and in long
    double, and one whose two evaluations differ by more than 1e-13
    (absolute, or relative to its largest integral above 1) is digested from
    its quad-precision integrals.  all_quad: every quartet in quad precision
    (slow; for small molecules).  Returns (G, J, K, n_quartets,
    n_quartets_in_quad)."""
    ob = OBasis.of(basis)
    D = np.ascontiguousarray(D, np.float64)
    N = D.shape[0]
    G = np.zeros((N, N))
    JK = np.zeros((2, N, N))
    nq = ctypes.c_int64()
    narb = ctypes.c_int64()
    L = lib()
    L.oracle_set_tight_centre(1)
    try:
        L.oracle_fock_2e_arbitrated(*ob.args(), D.reshape(-1), precision, G.reshape(-1), screening, JK.reshape(-1),
                                    nthreads, 2 if all_quad else 1, ctypes.byref(nq), ctypes.byref(narb))
    finally:
        L.oracle_set_tight_centre(0)
    return G, JK[0], JK[1], nq.value, narb.value


def fock_rows_exact(basis, D, shells, precision=1e-12, screening=SCHWARZ_INF, nthreads=0):
    """Rows of fock_2e_exact's G, J, K for the functions of `shells`, from
    only the quartets that touch those shells (all that add to their rows).
    Returns (rows, G_rows, J_rows, K_rows, n_quartets, n_quartets_in_quad)."""
    rows = shell_rows(basis, shells)
    with fock_rows_only(basis, shells):
        G, J, K, nq, nquad = fock_2e_exact(basis, D, precision, screening, nthreads)
    return rows, G[rows], J[rows], K[rows], nq, nquad


def shell_rows(basis, shells):
    """Basis-function indices of `shells`, in order."""
    ob = OBasis.of(basis)
    sizes = np.where(ob.pure != 0, 2 * ob.l + 1, (ob.l + 1) * (ob.l + 2) // 2)
    first = np.concatenate([[0], np.cumsum(sizes)])
    return np.concatenate([np.arange(first[s], first[s + 1]) for s in shells])


@contextlib.contextmanager
def fock_rows_only(basis, shells):
    """Within the block, the Fock builds digest only the quartets touching
    `shells`: the rows (and columns) of their functions come out whole, the
    rest of G, J, K is partial."""
    mask = np.zeros(len(OBasis.of(basis).l), np.int8)
    mask[np.asarray(shells, np.int64)] = 1
    L = lib()
    L.oracle_set_fock_rows(len(mask), mask.ctypes.data)
    try:
        yield
    finally:
        L.oracle_set_fock_rows(0, None)


def fock_pairs(basis, threshold=1e-12, tight_centre=False):
    """(significant-pair mask (s1 >= s2), Schwarz matrix)."""
    ob = OBasis.of(basis)
    n = len(ob.l)
    mask = np.zeros((n, n), np.int8)
    K = np.zeros((n, n))
    L = lib()
    L.oracle_set_tight_centre(int(bool(tight_centre)))
    try:
        L.oracle_fock_pairs(*ob.args(), threshold, mask.reshape(-1), K.reshape(-1))
    finally:
        L.oracle_set_tight_centre(0)
    return mask, K
