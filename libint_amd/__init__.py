"""libint_amd — MI355X (gfx950) evaluator for libint's two-electron Coulomb path.

Python host bindings over the C ABI in include/l2g.h (libint_amd/lib/libl2g.so).
The GPU work is done by the generated HIP kernels inside that library; this
module only moves arrays across the boundary.  There is no CPU fallback: if
the library is missing or no GPU is visible, the calls raise.

The objects mirror libint's C++ API for this path:

* :class:`Shell` — ``libint2::Shell`` (include/libint2/shell.h:720-1012):
  exponents, one contraction ``(l, pure, coeff)``, origin; coefficients are
  renormalised on construction (``Shell::renorm``, shell.h:958-999).
* :class:`Engine` — ``libint2::Engine(Operator::coulomb, max_nprim, max_l)``
  (include/libint2/engine.h:502-526) with ``compute(s1, s2, s3, s4)``
  returning the libint-ordered buffer, or ``None`` where libint returns
  ``results()[0] == nullptr`` (engine.impl.h:1781-1784), plus the batched
  ``compute_batch``.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import subprocess
import weakref
from typing import Iterable, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("LIBINT_AMD_LIB") or os.path.join(_HERE, "lib", "libl2g.so")

SCREEN_ORIGINAL, SCREEN_CONSERVATIVE, SCREEN_SCHWARZ, SCREEN_SCHWARZINF = 0x0001, 0x0010, 0x0100, 0x1000
SHG_STANDARD, SHG_GAUSSIAN = 1, 2
# error codes of include/l2g.h
L2G_OK, L2G_E_INVALID, L2G_E_LMAX, L2G_E_NODEVICE, L2G_E_OOM, L2G_E_LAUNCH, L2G_E_STATE = 0, -1, -2, -3, -4, -5, -6
REQ_QUEUED, REQ_ACTIVE, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p

# name -> (restype, argtypes) of every symbol in include/l2g.h
ABI = {
    "l2g_version": (ctypes.c_char_p, []),
    "l2g_last_error": (ctypes.c_char_p, []),
    "l2g_max_l": (ctypes.c_int, []),
    "l2g_fp32_max_l": (ctypes.c_int, []),
    "l2g_init": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "l2g_finalize": (ctypes.c_int, [_vp]),
    "l2g_set_solid_harmonics_ordering": (ctypes.c_int, [_vp, ctypes.c_int]),
    "l2g_shell_renorm": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _f64p, _f64p, _vp, ctypes.c_int]),
    "l2g_solid_harmonics_coeff": (ctypes.c_double, [ctypes.c_int] * 5),
    "l2g_boys_table": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "l2g_basis_create": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p,
                                        ctypes.POINTER(_vp)]),
    "l2g_basis_destroy": (ctypes.c_int, [_vp]),
    "l2g_eri_plan_create": (ctypes.c_int, [_vp, ctypes.c_int64, _i32p, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)]),
    "l2g_eri_plan_size": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "l2g_eri_plan_offsets": (ctypes.c_int, [_vp, _i64p]),
    "l2g_eri_plan_execute": (ctypes.c_int, [_vp, _vp]),
    "l2g_synchronize": (ctypes.c_int, [_vp]),
    "l2g_eri_plan_download": (ctypes.c_int, [_vp, _vp, _vp]),
    "l2g_eri_plan_device_output": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "l2g_eri_plan_stats": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp, _vp, _vp, _vp]),
    "l2g_eri_plan_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "l2g_eri_plan_set_wave_split": (ctypes.c_int, [_vp, ctypes.c_int]),
    "l2g_eri_plan_set_operator": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
    "l2g_eri_plan_elapsed": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "l2g_eri_plan_destroy": (ctypes.c_int, [_vp]),
    "l2g_onebody": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "l2g_eri_deriv1_plan_create": (ctypes.c_int, [_vp, ctypes.c_int64, _i32p, ctypes.c_double, ctypes.c_int,
                                                  ctypes.POINTER(_vp)]),
    "l2g_eri_deriv1_plan_size": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "l2g_eri_deriv1_plan_offsets": (ctypes.c_int, [_vp, _i64p]),
    "l2g_eri_deriv1_plan_execute": (ctypes.c_int, [_vp, _vp]),
    "l2g_eri_deriv1_plan_download": (ctypes.c_int, [_vp, _vp, _vp]),
    "l2g_eri_deriv1_plan_destroy": (ctypes.c_int, [_vp]),
    "l2g_eri_compute": (ctypes.c_int, [_vp, ctypes.c_int64, _i32p, ctypes.c_double, ctypes.c_int, _vp, _vp, _vp]),
    "l2g_schwarz": (ctypes.c_int, [_vp, ctypes.c_int64, _i32p, ctypes.c_double, ctypes.c_int, _f64p]),
    "l2g_fock_create": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)]),
    "l2g_fock_nbf": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "l2g_fock_schwarz": (ctypes.c_int, [_vp, _f64p]),
    "l2g_fock_build": (ctypes.c_int, [_vp, _f64p, ctypes.c_double, _vp, _vp, _vp]),
    "l2g_fock_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_float)]),
    "l2g_fock_destroy": (ctypes.c_int, [_vp]),
    "l2g_eri_plan_create_mixed": (ctypes.c_int, [_vp, ctypes.c_int64, _i32p, ctypes.c_double, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_vp)]),
    "l2g_fock_set_fp32_threshold": (ctypes.c_int, [_vp, ctypes.c_double]),
    "l2g_fock_stats_mixed": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "l2g_precision_bounds": (ctypes.c_int, [_vp, ctypes.c_int64, _i32p, _f64p, _f64p]),
    "l2g_fock_class_stats": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp, _vp, _vp, _vp]),
    "l2g_set_generic_kernels": (ctypes.c_int, [_vp, ctypes.c_int]),
    "l2g_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_vp)]),
    "l2g_request_submit": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                          ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int64)]),
    "l2g_request_state": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int)]),
    "l2g_request_wait": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "l2g_request_result": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp]),
    "l2g_server_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64)]),
    "l2g_server_destroy": (ctypes.c_int, [_vp]),
    "l2g_fock_set_window": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "l2g_fock_set_deterministic": (ctypes.c_int, [_vp, ctypes.c_int]),
    "l2g_set_fp32_policy": (ctypes.c_int, [_vp, ctypes.c_int]),
    "l2g_fp32_paying_classes": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp]),
    "l2g_fock_set_progress": (ctypes.c_int, [_vp, _vp, _vp]),
    "l2g_fock_count": (ctypes.c_int, [_vp, _f64p, ctypes.c_double, ctypes.POINTER(ctypes.c_int64)]),
    "l2g_fock_stats_windows": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "l2g_quartet_engine_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "l2g_quartet_engine_destroy": (ctypes.c_int, [_vp]),
    "l2g_quartet_compute": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p, _vp, _vp, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           _f64p, _vp]),
    "l2g_quartet_compute_noalloc": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p, _vp, _vp,
                                                   ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_double, _f64p, _vp]),
}


_PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p)
FP32_MEASURED, FP32_ALWAYS = 0, 1  # L2G_FP32_* of include/l2g.h


class ServerConfig(ctypes.Structure):
    """l2g_server_config of include/l2g.h."""
    _fields_ = [("device", ctypes.c_int), ("hbm_bytes", ctypes.c_double), ("max_active", ctypes.c_int),
                ("round_quartets", ctypes.c_int64), ("screening", ctypes.c_int), ("pair_threshold", ctypes.c_double),
                ("max_engine_precision", ctypes.c_double), ("deterministic", ctypes.c_int)]

_lib = None

# Every Python object that owns a native handle (servers, plans, Fock
# builders, bases, contexts).  At interpreter exit _release_all() frees them
# in dependency order while the library and the HIP runtime are still whole;
# finalizers that run after it find nothing left to free.
_live = weakref.WeakSet()
_exiting = False


def _release_all() -> None:
    """Frees every live native handle: servers (their threads use the device)
    first, then plans and Fock builders, then bases, then contexts.  Runs
    once, from atexit, before the interpreter tears its modules down."""
    global _exiting
    if _exiting:
        return
    objs = list(_live)
    rank = {Server: 0, Basis: 2, Context: 3}
    for o in sorted(objs, key=lambda o: rank.get(type(o), 1)):
        try:
            if isinstance(o, (Server, Context)):
                o.close()
            else:
                o._release()
        except Exception:
            pass
    _exiting = True


atexit.register(_release_all)


class L2gError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[l2g {code}] {msg}")
        self.code = code


class LmaxExceeded(L2gError):
    """libint2::Engine::lmax_exceeded (include/libint2/engine.h:893-916)."""


def build() -> None:
    """Generate and compile the gfx950 kernels and the C ABI library."""
    jobs = str(min(16, os.cpu_count() or 8))
    subprocess.run(["make", "-s", "-C", _HERE, "-j", jobs], check=True)


def lib():
    """Load libl2g.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} not built; run libint_amd.build() (or __graft_entry__.build())")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in ABI.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().l2g_last_error().decode(errors="replace")
        if rc == -2:
            raise LmaxExceeded(rc, msg)
        raise L2gError(rc, msg)


def max_l() -> int:
    return lib().l2g_max_l()


def fp32_max_l() -> int:
    """Largest l with FP32 kernel variants (mixed precision runs quartets
    with a higher l in FP64)."""
    return lib().l2g_fp32_max_l()


def fp32_paying_classes() -> list:
    """Classes ('2021', ...) whose FP32 variant was measured faster than
    their FP64 kernel: the ones mixed batches route to FP32 by default."""
    n = ctypes.c_int()
    codes = np.zeros(256, np.int32)
    _check(lib().l2g_fp32_paying_classes(256, ctypes.byref(n), codes.ctypes.data))
    return [f"{c:04d}" for c in codes[:min(n.value, 256)]]


def shell_renorm(l: int, alpha, coeff, enforce_unit: bool = True):
    a = np.ascontiguousarray(alpha, np.float64)
    c = np.array(coeff, np.float64)
    mlc = np.zeros_like(c)
    _check(lib().l2g_shell_renorm(l, a.size, a, c, mlc.ctypes.data, 1 if enforce_unit else 0))
    return c, mlc


def solid_harmonics_coeff(l, m, lx, ly, lz) -> float:
    return lib().l2g_solid_harmonics_coeff(l, m, lx, ly, lz)


def boys_table() -> np.ndarray:
    """Host copy of the Boys table, shape (mmax+1, nint, 8)."""
    nint, mmax = ctypes.c_int(), ctypes.c_int()
    _check(lib().l2g_boys_table(None, ctypes.byref(nint), ctypes.byref(mmax)))
    out = np.zeros((mmax.value + 1, nint.value, 8))
    _check(lib().l2g_boys_table(out.ctypes.data, ctypes.byref(nint), ctypes.byref(mmax)))
    return out


# ---------------------------------------------------------------------------
# shells and basis arrays
# ---------------------------------------------------------------------------


class Shell:
    """Single-contraction Gaussian shell, coefficients renormalised as
    libint2::Shell does (shell.h:767-777 with renorm :958-999).

    ``Shell(l, [(alpha, coeff), ...], origin=(0,0,0), pure=...)`` mirrors the
    python binding's ``Shell(L, primitives, center, pure)``
    (python/src/libint2/libint2.cc:155-156)."""

    def __init__(self, l: int, primitives: Sequence, origin=(0.0, 0.0, 0.0), pure: bool = True,
                 renorm: bool = True):
        self.l = int(l)
        self.pure = bool(pure)
        self.alpha = np.array([p[0] for p in primitives], np.float64)
        c = np.array([p[1] for p in primitives], np.float64)
        if renorm:
            c, _ = shell_renorm(self.l, self.alpha, c)
        self.coeff = c
        self.O = np.array(origin, np.float64).reshape(3)

    @staticmethod
    def unit() -> "Shell":
        """libint2::Shell::unit() (shell.h:904-909): l = 0, exponent 0,
        coefficient 1 (not renormalised), at the origin; stands for the
        missing centre of 3- and 2-centre integrals."""
        return Shell(0, [(0.0, 1.0)], pure=False, renorm=False)

    def size(self) -> int:
        return 2 * self.l + 1 if self.pure else (self.l + 1) * (self.l + 2) // 2

    def nprim(self) -> int:
        return self.alpha.size


class BasisArrays:
    """Flat per-shell / per-primitive arrays, the form the C ABI takes."""

    def __init__(self, l, pure, nprim, alpha, coeff, origin):
        self.l = np.ascontiguousarray(l, np.int32)
        self.pure = np.ascontiguousarray(pure, np.int32)
        self.nprim = np.ascontiguousarray(nprim, np.int32)
        self.alpha = np.ascontiguousarray(alpha, np.float64)
        self.coeff = np.ascontiguousarray(coeff, np.float64)
        self.origin = np.ascontiguousarray(np.asarray(origin, np.float64).reshape(-1))

    @staticmethod
    def from_shells(shells: Iterable[Shell]) -> "BasisArrays":
        shells = list(shells)
        return BasisArrays([s.l for s in shells], [int(s.pure) for s in shells], [s.nprim() for s in shells],
                           np.concatenate([s.alpha for s in shells]) if shells else np.zeros(0),
                           np.concatenate([s.coeff for s in shells]) if shells else np.zeros(0),
                           np.concatenate([s.O for s in shells]) if shells else np.zeros(0))

    @property
    def nshell(self) -> int:
        return int(self.l.size)

    def sizes(self) -> np.ndarray:
        return np.where(self.pure != 0, 2 * self.l + 1, (self.l + 1) * (self.l + 2) // 2)


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------


class Context:
    """HIP device context (libint2::initialize for this path)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().l2g_init(device, ctypes.byref(h)))
        self._h = h
        # live bases, plans and Fock builders: released before the context
        self._children = weakref.WeakSet()
        _live.add(self)

    @property
    def handle(self):
        return self._h

    def set_solid_harmonics_ordering(self, ordering: int) -> None:
        _check(lib().l2g_set_solid_harmonics_ordering(self._h, ordering))

    def set_fp32_policy(self, policy: int) -> None:
        """FP32_MEASURED (default): mixed batches use the FP32 variants of
        the classes where they were measured faster; FP32_ALWAYS: of every
        class that has one (l2g_set_fp32_policy)."""
        _check(lib().l2g_set_fp32_policy(self._h, int(policy)))

    def set_generic_kernels(self, on: bool) -> None:
        """Route every class through the generic kernel (a testing aid)."""
        _check(lib().l2g_set_generic_kernels(self._h, int(on)))

    def synchronize(self) -> None:
        _check(lib().l2g_synchronize(self._h))

    def close(self) -> None:
        """Releases every object still alive on this context (plans and Fock
        builders first, then bases), then the context itself."""
        if self._h:
            kids = list(self._children)
            for k in kids:
                if not isinstance(k, Basis):
                    k._release()
            for k in kids:
                if isinstance(k, Basis):
                    k._release()
            _check(lib().l2g_finalize(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Basis:
    def __init__(self, ctx: Context, arrays: BasisArrays):
        self.ctx = ctx
        self.arrays = arrays
        h = _vp()
        a = arrays
        _check(lib().l2g_basis_create(ctx.handle, a.nshell, a.l, a.pure, a.nprim, a.alpha, a.coeff, a.origin,
                                      ctypes.byref(h)))
        self._h = h
        ctx._children.add(self)
        _live.add(self)

    @property
    def handle(self):
        if not self._h:
            raise L2gError(L2G_E_STATE, "basis released (its context was closed)")
        return self._h

    def _release(self):
        if self._h:
            lib().l2g_basis_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


class EriPlan:
    """A planned batch of shell quartets (caller order)."""

    def __init__(self, basis: Basis, quartets, precision: float = np.finfo(float).eps,
                 screening: int = SCREEN_ORIGINAL, scale: float = 1.0, uniform_cart: bool = False,
                 fp32_threshold: float = 0.0):
        self.basis = basis
        q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
        self.nq = q.shape[0]
        h = _vp()
        if fp32_threshold > 0:
            _check(lib().l2g_eri_plan_create_mixed(basis.handle, self.nq, q.reshape(-1), precision, screening,
                                                   scale, int(uniform_cart), fp32_threshold, ctypes.byref(h)))
        else:
            _check(lib().l2g_eri_plan_create(basis.handle, self.nq, q.reshape(-1), precision, screening, scale,
                                             int(uniform_cart), ctypes.byref(h)))
        self._h = h
        basis.ctx._children.add(self)
        _live.add(self)
        n = ctypes.c_int64()
        _check(lib().l2g_eri_plan_size(h, ctypes.byref(n)))
        self.size = n.value

    def offsets(self) -> np.ndarray:
        o = np.zeros(self.nq, np.int64)
        if self.nq:
            _check(lib().l2g_eri_plan_offsets(self._h, o))
        return o

    def execute(self, d_out: int | None = None) -> None:
        _check(lib().l2g_eri_plan_execute(self._h, d_out))

    def device_output(self) -> int:
        p = _vp()
        _check(lib().l2g_eri_plan_device_output(self._h, ctypes.byref(p)))
        return p.value

    def download(self):
        out = np.zeros(max(self.size, 1))
        nz = np.zeros(max(self.nq, 1), np.int8)
        _check(lib().l2g_eri_plan_download(self._h, out.ctypes.data, nz.ctypes.data))
        return out[: self.size], nz[: self.nq]

    def stats(self):
        cap = 256
        n = ctypes.c_int()
        code = np.zeros(cap, np.int32)
        nq = np.zeros(cap, np.int64)
        npq = np.zeros(cap, np.int64)
        ms = np.zeros(cap, np.float32)
        _check(lib().l2g_eri_plan_stats(self._h, cap, ctypes.byref(n), code.ctypes.data, nq.ctypes.data,
                                        npq.ctypes.data, ms.ctypes.data))
        k = n.value
        return [dict(cls=f"{c % 10000:04d}", fp32=bool(c % 100000 // 10000 == 1),
                     generic=bool(c % 100000 // 10000 == 2), wave=bool(c >= 100000),
                     nquartets=int(a), nprim=int(b), ms=float(t))
                for c, a, b, t in zip(code[:k], nq[:k], npq[:k], ms[:k])]

    def set_graph(self, on=True):
        """Capture the class launches into a hipGraph on the next execution
        and replay it on later executions into the same output buffer."""
        _check(lib().l2g_eri_plan_set_graph(self._h, int(bool(on))))

    def set_operator(self, oper: int, omega: float = 0.0):
        """COULOMB, ERF_COULOMB or ERFC_COULOMB with attenuation omega
        (libint's Operator::coulomb / erf_coulomb / erfc_coulomb)."""
        _check(lib().l2g_eri_plan_set_operator(self._h, int(oper), float(omega)))

    def set_wave_split(self, on=True):
        """Classes of few, deeply contracted quartets run a wave per quartet
        (lanes split the primitive loop; on by default): each contraction is
        summed as 64 partial sums and a fixed butterfly instead of libint's
        serial order.  True (1) routes the classes measured faster on the
        generic kernel there instead; 2 keeps them all on their register
        kernels; False (0) turns the split off."""
        _check(lib().l2g_eri_plan_set_wave_split(self._h, int(on)))

    def elapsed_ms(self):
        """Duration of the last execution (all classes), waiting for it."""
        t = ctypes.c_float()
        _check(lib().l2g_eri_plan_elapsed(self._h, ctypes.byref(t)))
        return float(t.value)

    def _release(self):
        if self._h:
            lib().l2g_eri_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


def one_body(basis: Basis, charges):
    """Overlap, kinetic energy and point-charge attraction matrices (S, T, V)
    over the basis's functions; charges: (Z, (x, y, z)) pairs in bohr, as
    the reference's compute_1body_ints gives the nuclei to
    Engine(Operator::nuclear) (tests/hartree-fock/hartree-fock.cc:480-535)."""
    ch = list(charges)
    q = np.ascontiguousarray([float(z) for z, _ in ch], np.float64)
    xyz = np.ascontiguousarray([np.asarray(x, np.float64).reshape(3) for _, x in ch], np.float64).reshape(-1)
    n = int(basis.arrays.sizes().sum()) if basis.arrays.nshell else 0
    S, T, V = (np.zeros((max(n, 1), max(n, 1))) for _ in range(3))
    _check(lib().l2g_onebody(basis.handle, len(ch), q.ctypes.data if len(ch) else None,
                             xyz.ctypes.data if len(ch) else None, S.ctypes.data, T.ctypes.data,
                             V.ctypes.data if len(ch) else None))
    return S[:n, :n], T[:n, :n], V[:n, :n]


class Deriv1Plan:
    """First geometric derivatives of a batch of quartets (deriv_order = 1):
    12 shell sets per quartet, d = 3 * centre + axis in caller order."""

    def __init__(self, basis: Basis, quartets, precision: float = np.finfo(float).eps,
                 screening: int = SCREEN_ORIGINAL):
        self.basis = basis
        q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
        self.nq = q.shape[0]
        h = _vp()
        _check(lib().l2g_eri_deriv1_plan_create(basis.handle, self.nq, q.reshape(-1), precision, screening,
                                                ctypes.byref(h)))
        self._h = h
        basis.ctx._children.add(self)
        _live.add(self)
        n = ctypes.c_int64()
        _check(lib().l2g_eri_deriv1_plan_size(h, ctypes.byref(n)))
        self.size = n.value

    def offsets(self) -> np.ndarray:
        o = np.zeros(self.nq, np.int64)
        if self.nq:
            _check(lib().l2g_eri_deriv1_plan_offsets(self._h, o))
        return o

    def execute(self, d_out: int | None = None) -> None:
        _check(lib().l2g_eri_deriv1_plan_execute(self._h, d_out))

    def download(self):
        """(values of shape (12, size), nonzero[nq])"""
        out = np.zeros(max(12 * self.size, 1))
        nz = np.zeros(max(self.nq, 1), np.int8)
        _check(lib().l2g_eri_deriv1_plan_download(self._h, out.ctypes.data, nz.ctypes.data))
        return out[: 12 * self.size].reshape(12, self.size), nz[: self.nq]

    def _release(self):
        if self._h:
            lib().l2g_eri_deriv1_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


def schwarz(basis: Basis, pairs, precision: float = 0.0, norm: str = "inf") -> np.ndarray:
    """Schwarz factors for shell pairs (n, 2): sqrt(max|(ab|ab)|) ("inf") or
    sqrt(||(ab|ab)||_2) ("2"), as compute_schwarz_ints
    (tests/hartree-fock/hartree-fock++.cc:1231-1298)."""
    p = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
    K = np.zeros(max(p.shape[0], 1))
    _check(lib().l2g_schwarz(basis.handle, p.shape[0], p.reshape(-1), precision, 0 if norm == "inf" else 1, K))
    return K[: p.shape[0]]


def precision_bounds(basis: Basis, pairs):
    """(K, S) per shell pair (n, 2): the Schwarz factor and the pre-HRR
    factor that together select FP32 in mixed batches; a quartet runs in FP32
    when K_ab K_cd < threshold and S_ab S_cd < 16 threshold."""
    p = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
    K = np.zeros(max(p.shape[0], 1))
    S = np.zeros(max(p.shape[0], 1))
    _check(lib().l2g_precision_bounds(basis.handle, p.shape[0], p.reshape(-1), K, S))
    return K[: p.shape[0]], S[: p.shape[0]]


FP32_HRR_FACTOR = 16.0

# Operator of an ERI plan (l2g.h L2G_OPER_*; libint's Operator::coulomb,
# erf_coulomb, erfc_coulomb, engine.h:158-167)
COULOMB, ERF_COULOMB, ERFC_COULOMB = 0, 1, 2
# one-body operators of the Engine (Operator::overlap / kinetic / nuclear), BraKet "XX"
OVERLAP, KINETIC, NUCLEAR = "overlap", "kinetic", "nuclear"
ONE_BODY = (OVERLAP, KINETIC, NUCLEAR)


class FockBuilder:
    """Screened two-electron Fock builds on the GPU for one basis
    (compute_shellpairs + compute_schwarz_ints + compute_2body_fock of
    tests/hartree-fock/hartree-fock++.cc)."""

    def __init__(self, basis: Basis, pair_threshold: float = 1e-12,
                 max_engine_precision: float = np.finfo(float).eps / 1e10, screening: int = SCREEN_SCHWARZINF):
        self.basis = basis
        h = _vp()
        _check(lib().l2g_fock_create(basis.handle, pair_threshold, max_engine_precision, screening, ctypes.byref(h)))
        self._h = h
        basis.ctx._children.add(self)
        _live.add(self)
        n = ctypes.c_int()
        _check(lib().l2g_fock_nbf(h, ctypes.byref(n)))
        self.nbf = n.value

    def schwarz(self) -> np.ndarray:
        n = self.basis.arrays.nshell
        K = np.zeros(n * n)
        _check(lib().l2g_fock_schwarz(self._h, K))
        return K.reshape(n, n)

    def build(self, D, precision: float = 1e-12):
        """Returns (G, J, K) with G = 2J - K."""
        D = np.ascontiguousarray(D, np.float64)
        N = self.nbf
        assert D.shape == (N, N)
        G, J, K = np.zeros((N, N)), np.zeros((N, N)), np.zeros((N, N))
        _check(lib().l2g_fock_build(self._h, D.reshape(-1), precision, J.ctypes.data, K.ctypes.data, G.ctypes.data))
        return G, J, K

    def set_fp32_threshold(self, threshold: float) -> None:
        """Quartets with Schwarz bound below `threshold` run in FP32."""
        _check(lib().l2g_fock_set_fp32_threshold(self._h, threshold))

    def set_progress(self, fn) -> None:
        """fn(done, total) after each window of a build (None: off)."""
        self._progress = None if fn is None else _PROGRESS_FN(lambda d, tot, u: fn(int(d), int(tot)))
        ptr = None if self._progress is None else ctypes.cast(self._progress, ctypes.c_void_p)
        _check(lib().l2g_fock_set_progress(self._h, ptr, None))

    def set_deterministic(self, on: bool = True) -> None:
        """Fixed-point J/K accumulation: the same bits on every run
        (l2g_fock_set_deterministic)."""
        _check(lib().l2g_fock_set_deterministic(self._h, int(bool(on))))

    def count(self, D, precision: float = 1e-12) -> int:
        """Screened unique quartets a build of D at this precision digests,
        from the build's count pass alone (l2g_fock_count)."""
        D = np.ascontiguousarray(D, np.float64)
        assert D.shape == (self.nbf, self.nbf)
        n = ctypes.c_int64()
        _check(lib().l2g_fock_count(self._h, D.reshape(-1), precision, ctypes.byref(n)))
        return n.value

    def set_window(self, items: int) -> None:
        """Quartets per window of the build's list (0: the default)."""
        _check(lib().l2g_fock_set_window(self._h, int(items)))

    def stats(self):
        a, b, c, d, w = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_float(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().l2g_fock_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        _check(lib().l2g_fock_stats_mixed(self._h, ctypes.byref(d)))
        _check(lib().l2g_fock_stats_windows(self._h, ctypes.byref(w)))
        return dict(nquartets=a.value, nprim=b.value, kernel_ms=c.value, nquartets_fp32=d.value, windows=w.value)

    def class_stats(self):
        """Per kernel class of the last build: class, fp32, quartets,
        primitive quartets and kernel ms."""
        cap = 256
        n = ctypes.c_int()
        code = np.zeros(cap, np.int32)
        nq = np.zeros(cap, np.int64)
        npq = np.zeros(cap, np.int64)
        ms = np.zeros(cap, np.float32)
        _check(lib().l2g_fock_class_stats(self._h, cap, ctypes.byref(n), code.ctypes.data, nq.ctypes.data,
                                          npq.ctypes.data, ms.ctypes.data))
        k = n.value
        return [dict(cls=f"{c % 10000:04d}", fp32=bool(c // 10000 == 1), generic=bool(c // 10000 == 2),
                     nquartets=int(a), nprim=int(b), ms=float(t))
                for c, a, b, t in zip(code[:k], nq[:k], npq[:k], ms[:k])]

    def _release(self):
        if self._h:
            lib().l2g_fock_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


class QuartetEngine:
    """Resident single-quartet evaluator (l2g_quartet_engine_* of
    include/l2g.h): device buffers and pinned staging are allocated once, and
    each compute() makes one copy in, one launch per operator pass and one
    copy out.  The per-call path of Engine.compute."""

    def __init__(self, ctx: Context, max_nprim: int, max_l: int):
        h = _vp()
        _check(lib().l2g_quartet_engine_create(ctx.handle, max_nprim, max_l, ctypes.byref(h)))
        self._h = h
        self.ctx = ctx
        ctx._children.add(self)
        _live.add(self)

    def compute(self, shells, precision=np.finfo(float).eps, screening=SCREEN_ORIGINAL, scale=1.0,
                uniform_cart=False, oper=0, omega=0.0, ln_scr=None, nonsph=None):
        """Integrals of (s0 s1|s2 s3) in libint's layout and the nonzero
        flag.  ln_scr / nonsph: the precomputed shell pairs' primitive-pair
        screening data (bra n0 n1 values, then ket n2 n3; -inf drops a
        primitive pair), as compute2 with ShellPair arguments."""
        a = BasisArrays.from_shells(shells)
        n = int(np.prod(a.sizes()))
        out = np.zeros(max(n, 1))
        nz = np.zeros(1, np.int8)
        ls = None if ln_scr is None else np.ascontiguousarray(ln_scr, np.float64)
        ns = None if nonsph is None else np.ascontiguousarray(nonsph, np.float64)
        _check(lib().l2g_quartet_compute(self._h, a.l, a.pure, a.nprim, a.alpha, a.coeff, a.origin,
                                         None if ls is None else ls.ctypes.data,
                                         None if ns is None else ns.ctypes.data, precision, screening, scale,
                                         int(uniform_cart), oper, omega, out, nz.ctypes.data))
        return out[:n], bool(nz[0])

    def _release(self):
        if self._h:
            lib().l2g_quartet_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


NOALLOC_MAX_NPRIM = 16  # L2G_NOALLOC_MAX_NPRIM of include/l2g.h


def quartet_compute_noalloc(ctx: Context, shells, precision=np.finfo(float).eps, screening=SCREEN_ORIGINAL,
                            scale=1.0, uniform_cart=False, oper=0, omega=0.0, ln_scr=None, nonsph=None):
    """QuartetEngine.compute without an engine: l2g_quartet_compute_noalloc,
    which allocates no device or pinned memory (the library's fixed scratch
    array of the context's device; l <= 3, at most NOALLOC_MAX_NPRIM
    primitives per shell).  Returns the integrals and the nonzero flag."""
    a = BasisArrays.from_shells(shells)
    n = int(np.prod(a.sizes()))
    out = np.zeros(max(n, 1))
    nz = np.zeros(1, np.int8)
    ls = None if ln_scr is None else np.ascontiguousarray(ln_scr, np.float64)
    ns = None if nonsph is None else np.ascontiguousarray(nonsph, np.float64)
    _check(lib().l2g_quartet_compute_noalloc(ctx.handle, a.l, a.pure, a.nprim, a.alpha, a.coeff, a.origin,
                                             None if ls is None else ls.ctypes.data,
                                             None if ns is None else ns.ctypes.data, precision, screening, scale,
                                             int(uniform_cart), oper, omega, out, nz.ctypes.data))
    return out[:n], bool(nz[0])


def compute_quartets(basis: Basis, quartets, precision=np.finfo(float).eps, screening=SCREEN_ORIGINAL,
                     scale=1.0, uniform_cart=False, fp32_threshold=0.0):
    """Evaluate a batch; returns (values, offsets, nonzero)."""
    p = EriPlan(basis, quartets, precision, screening, scale, uniform_cart, fp32_threshold)
    offs = p.offsets()
    p.execute()
    vals, nz = p.download()
    return vals, offs, nz


def uniform_cart_scales(shell: "Shell") -> np.ndarray:
    """Per-function factors that take a shell from libint's standard to its
    uniform Cartesian normalization: sqrt((2l-1)!! / ((2i-1)!! (2j-1)!!
    (2k-1)!!)) for Cartesian component (i, j, k), 1 for a solid-harmonic shell
    (include/libint2/cartesian.h:96-133)."""
    if shell.pure:
        return np.ones(2 * shell.l + 1)

    def dfact(k):
        r = 1.0
        for i in range(2 * k - 1, 1, -2):
            r *= i
        return r

    l = shell.l
    comps = [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]
    return np.array([np.sqrt(dfact(l) / (dfact(x) * dfact(y) * dfact(z))) for x, y, z in comps])


_default_ctx = None


def default_context() -> "Context":
    """The process's shared context on device 0, opened on first use (the
    Engine's default when no context is given); released at exit."""
    global _default_ctx
    if _default_ctx is None or not _default_ctx._h:
        _default_ctx = Context(0)
    return _default_ctx


class Engine:
    """Coulomb ``libint2::Engine`` for this path (engine.h:502-526).

    ``compute(s1, s2, s3, s4)`` evaluates one quartet on the GPU and returns
    a flat array in libint's layout, or None where libint returns nullptr.
    One Engine per thread, as libint requires (engine.h:560-579)."""

    # BraKet of the reference's python binding (python/src/libint2/libint2.cc:212-217):
    # where compute()'s shells go in the 4-centre quartet; None = Shell.unit()
    BRAKETS = {"XXXX": (0, 1, 2, 3), "XSXX": (0, None, 1, 2), "XXXS": (0, 1, 2, None),
               "XSXS": (0, None, 1, None)}

    def __init__(self, max_nprim: int, max_l: int, precision: float = np.finfo(float).eps,
                 screening: int = SCREEN_ORIGINAL, ctx: Context | None = None, braket: str | None = None,
                 deriv_order: int = 0, oper=None):
        oper = COULOMB if oper is None else oper
        if oper not in (COULOMB, ERF_COULOMB, ERFC_COULOMB) + ONE_BODY:
            raise ValueError("oper must be COULOMB, ERF_COULOMB, ERFC_COULOMB, OVERLAP, KINETIC or NUCLEAR")
        braket = braket or ("XX" if oper in ONE_BODY else "XXXX")
        if (braket == "XX") != (oper in ONE_BODY) or (braket != "XX" and braket not in self.BRAKETS):
            raise ValueError(f"braket must be XX for one-body operators, else one of {sorted(self.BRAKETS)}")
        if deriv_order not in (0, 1) or (deriv_order and (braket != "XXXX" or oper != COULOMB)):
            raise ValueError("deriv_order 0, or 1 for 4-centre Coulomb integrals")
        self.oper = oper
        self.omega = 0.0
        self.charges = []
        self.braket = braket
        self.deriv_order = deriv_order
        if max_nprim <= 0:
            raise ValueError("max_nprim must be positive")
        if max_l > globals()["max_l"]():
            raise LmaxExceeded(-2, f"max_l {max_l} exceeds the compiled limit {globals()['max_l']()}")
        self.max_nprim = max_nprim
        self.max_l_ = max_l
        self.ctx = ctx or default_context()
        self.set_precision(precision)
        self.screening = screening
        self.scale = 1.0
        self.uniform_cart = False
        self._qe = None  # the resident single-quartet evaluator, made on first use

    def set_precision(self, prec: float) -> "Engine":
        self.precision = max(0.0, float(prec))
        return self

    def set_params(self, params) -> "Engine":
        """omega for ERF_COULOMB / ERFC_COULOMB; the point charges
        [(q, (x, y, z)), ...] for NUCLEAR (engine.h:685-690)."""
        if self.oper == NUCLEAR:
            self.charges = [(float(q), np.asarray(x, float).reshape(3)) for q, x in params]
        else:
            self.omega = float(params)
        return self

    def prescale_by(self, s: float) -> "Engine":
        self.scale = float(s)
        return self

    def set_cartesian_normalization(self, uniform: bool) -> "Engine":
        """CartesianShellNormalization::uniform (True) or ::standard (False),
        as Engine::set(CartesianShellNormalization) (engine.h:842-845)."""
        self.uniform_cart = bool(uniform)
        return self

    def compute(self, *shells: Shell):
        """compute(s1, s2, s3, s4) for XXXX; compute(a, c, d) = (a|cd) for
        XSXX and compute(a, b, c) = (ab|c) for XXXS (3-centre); compute(a, c)
        = (a|c) for XSXS (2-centre), the missing centres being Shell.unit()
        as Engine::compute does (engine.impl.h:140-178)."""
        if self.oper in ONE_BODY:
            if len(shells) != 2:
                raise ValueError("one-body operators take 2 shells")
            if self.oper == NUCLEAR and not self.charges:
                raise ValueError("NUCLEAR needs the charges (set_params)")
            for s_ in shells:
                if s_.l > self.max_l_:
                    raise LmaxExceeded(-2, "the angular momentum limit is exceeded")
            S, T, V = one_body(Basis(self.ctx, BasisArrays.from_shells(shells)),
                               self.charges if self.oper == NUCLEAR else [])
            M = {OVERLAP: S, KINETIC: T, NUCLEAR: V}[self.oper]
            n1 = shells[0].size()
            # prescale_by applies to one-body integrals too (engine.impl.h:1050);
            # uniform normalization rescales each Cartesian shell's functions
            # (engine.impl.h:552-558)
            blk = self.scale * np.array(M[:n1, n1:])
            if self.uniform_cart:
                blk *= np.outer(uniform_cart_scales(shells[0]), uniform_cart_scales(shells[1]))
            return np.ascontiguousarray(blk).reshape(-1)
        slots = self.BRAKETS[self.braket]
        if len(shells) != sum(k is not None for k in slots):
            raise ValueError(f"{self.braket} takes {sum(k is not None for k in slots)} shells")
        for s in shells:
            if s.l > self.max_l_:
                raise LmaxExceeded(-2, "the angular momentum limit is exceeded")
            if s.nprim() > self.max_nprim:
                raise ValueError("shell has more primitives than max_nprim")
        quad = [shells[k] if k is not None else Shell.unit() for k in slots]
        if self.deriv_order == 1:
            if self.scale != 1.0 or self.uniform_cart:
                raise ValueError("deriv_order 1 supports neither prescale_by nor uniform Cartesian normalization")
            # 12 shell sets, d/dA_x .. d/dD_z (libint's results()[0..11])
            from .deriv import eri_deriv1
            vals, _, nz = eri_deriv1(self.ctx, BasisArrays.from_shells(quad), [[0, 1, 2, 3]], self.precision,
                                     self.screening)
            return [vals[d] for d in range(12)] if nz[0] else None
        if self._qe is None:
            self._qe = QuartetEngine(self.ctx, self.max_nprim, self.max_l_)
        vals, nz = self._qe.compute(quad, self.precision, self.screening, self.scale, self.uniform_cart,
                                    self.oper, self.omega)
        return vals if nz else None

    def compute_batch(self, basis: Basis, quartets):
        return compute_quartets(basis, quartets, self.precision, self.screening, self.scale, self.uniform_cart)


class Server:
    """Fock-build request server (l2g_server_* of include/l2g.h): requests
    are admitted as slots and device memory free up, and their quartets share
    launch rounds."""

    def __init__(self, device: int = 0, hbm_bytes: float = 0.0, max_active: int = 0, round_quartets: int = 0,
                 screening: int = 0, pair_threshold: float = 0.0, max_engine_precision: float = 0.0,
                 deterministic: bool = False):
        cfg = ServerConfig(device, hbm_bytes, max_active, round_quartets, screening, pair_threshold,
                           max_engine_precision, int(bool(deterministic)))
        h = _vp()
        _check(lib().l2g_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._nbf = {}
        _live.add(self)

    def submit(self, arrays: BasisArrays, D, precision: float = 1e-12, fp32_threshold: float = 0.0) -> int:
        a = arrays
        D = np.ascontiguousarray(D, np.float64)
        nbf = int(a.sizes().sum())
        assert D.shape == (nbf, nbf)
        rid = ctypes.c_int64()
        _check(lib().l2g_request_submit(self._h, a.nshell, a.l, a.pure, a.nprim, a.alpha, a.coeff, a.origin,
                                        D.reshape(-1), precision, fp32_threshold, ctypes.byref(rid)))
        self._nbf[rid.value] = nbf
        return rid.value

    def state(self, rid: int) -> int:
        st = ctypes.c_int()
        _check(lib().l2g_request_state(self._h, rid, ctypes.byref(st)))
        return st.value

    def wait(self, rid: int) -> None:
        _check(lib().l2g_request_wait(self._h, rid))

    def result(self, rid: int):
        """Waits for the request and returns (G, J, K, ms) with ms = (submit to
        admission, submit to done) in milliseconds; releases the request."""
        self.wait(rid)
        n = self._nbf.pop(rid)
        G, J, K = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        ms = np.zeros(2)
        _check(lib().l2g_request_result(self._h, rid, J.ctypes.data, K.ctypes.data, G.ctypes.data, ms.ctypes.data))
        return G, J, K, ms

    def stats(self):
        r, m, q = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().l2g_server_stats(self._h, ctypes.byref(r), ctypes.byref(m), ctypes.byref(q)))
        return dict(rounds=r.value, max_requests_per_round=m.value, work_items=q.value)

    def close(self) -> None:
        if self._h:
            lib().l2g_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

