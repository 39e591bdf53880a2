"""ctypes binding of include/fastkpm_host.h (libfastkpm.so): the C++ host code
of the engine API -- std::mt19937 probing vectors (Engine<T>::set_R_*) and the
spectral functions.  Fails loudly when the library has not been built."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _capi

LIB_PATH = Path(__file__).resolve().parent / "libfastkpm.so"
_lib = None

_I, _D, _P, _U32 = C.c_int, C.c_double, C.c_void_p, C.c_uint32
_IP, _DP = C.POINTER(C.c_int), C.POINTER(C.c_double)
SIGNATURES = {
    "fkpm_host_last_error": (C.c_char_p, []),
    "fkpm_host_rng_new": (_P, [_U32]),
    "fkpm_host_rng_free": (None, [_P]),
    "fkpm_host_rng_next": (_U32, [_P]),
    "fkpm_host_rng_discard": (None, [_P, C.c_ulonglong]),
    "fkpm_host_set_R_identity": (_I, [_I, _I, _I, _I, _P]),
    "fkpm_host_set_R_uncorrelated": (_I, [_I, _I, _I, _P, _I, _I, _P]),
    "fkpm_host_set_R_correlated": (_I, [_I, _I, _IP, _P, _I, _I, _P, _IP]),
    "fkpm_host_set_R_probing": (_I, [_I, _IP, _P, _IP, _P, _IP]),
    "fkpm_host_expansion_coefficients": (_I, [_I, _I, _I, _D, _D, _D, _D, _DP]),
    "fkpm_host_conductivity_coefficients": (_I, [_I, _I, _D, _D, _D, _D, _D, _DP, _DP]),
    "fkpm_host_moment_transform2": (_I, [_I, _DP, _I, _DP, _DP]),
    "fkpm_host_jackson_kernel": (None, [_I, _DP]),
    "fkpm_host_moment_product": (_D, [_I, _DP, _DP]),
    "fkpm_host_moment_transform": (None, [_I, _DP, _I, _DP]),
    "fkpm_host_electronic_grand_energy": (_D, [_I, _DP, _D, _D, _D, _D]),
    "fkpm_host_filling_to_mu": (_D, [_I, _DP, _D, _D, _D, _D, _D]),
}

NP_OF = {_capi.F32: np.float32, _capi.F64: np.float64, _capi.C32: np.complex64,
         _capi.C64: np.complex128}
CODE_OF = {"float": _capi.F32, "double": _capi.F64, "cx_float": _capi.C32,
           "cx_double": _capi.C64}

FERMI_ENERGY, FERMI_DENSITY, CONSTANT_ONE = 0, 1, 2


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _capi.lib()  # the host library links the device library
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (fastkpm_amd/build.py)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class RNG:
    """fkpm::RNG (std::mt19937) living in the C++ library."""

    def __init__(self, seed: int = 5489):
        self._p = lib().fkpm_host_rng_new(seed & 0xFFFFFFFF)

    def __del__(self):
        try:
            lib().fkpm_host_rng_free(self._p)
        except Exception:
            pass

    def __call__(self) -> int:
        return int(lib().fkpm_host_rng_next(self._p))

    def discard(self, k: int) -> None:
        lib().fkpm_host_rng_discard(self._p, int(k))

    @property
    def ptr(self):
        return self._p


def _check(rc, what):
    if rc != 0:
        msg = lib().fkpm_host_last_error()
        raise ValueError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def set_R_identity(dtype: str, n: int, j0: int = 0, j1: int | None = None) -> np.ndarray:
    j1 = n if j1 is None else j1
    code = CODE_OF[dtype]
    R = np.zeros((n, j1 - j0), dtype=NP_OF[code], order="F")
    _check(lib().fkpm_host_set_R_identity(code, n, j0, j1, R.ctypes.data_as(_P)), "set_R_identity")
    return R


def set_R_uncorrelated(dtype: str, n: int, s: int, rng: RNG, j0: int = 0,
                       j1: int | None = None) -> np.ndarray:
    j1 = s if j1 is None else j1
    code = CODE_OF[dtype]
    R = np.zeros((n, j1 - j0), dtype=NP_OF[code], order="F")
    _check(lib().fkpm_host_set_R_uncorrelated(code, n, s, rng.ptr, j0, j1, R.ctypes.data_as(_P)),
           "set_R_uncorrelated")
    return R


def set_R_probing(groups, rng: RNG):
    """Compact form of the complex-double set_R_correlated(groups, rng): per
    orbital row the column of its one nonzero (-1: none) and its value.
    Returns (rcol int32[n], rval complex128[n], s)."""
    g = np.ascontiguousarray(groups, dtype=np.int32)
    n = len(g)
    rcol = np.empty(n, dtype=np.int32)
    rval = np.empty(n, dtype=np.complex128)
    s = C.c_int()
    _check(lib().fkpm_host_set_R_probing(n, g.ctypes.data_as(_IP), rng.ptr,
                                         rcol.ctypes.data_as(_IP), rval.ctypes.data_as(_P),
                                         C.byref(s)), "set_R_probing")
    return rcol, rval, s.value


def master_seeds(rng: RNG, s: int) -> np.ndarray:
    """The s draws of the caller's engine that set_R_correlated spends on
    seeding the columns' generators (src/engine.cpp:84-86), in order: the
    seeds a device-side generation of the same R takes."""
    return np.array([rng() for _ in range(int(s))], dtype=np.uint32)


def set_R_correlated(dtype: str, groups, rng: RNG, j0: int | None = None,
                     j1: int | None = None) -> np.ndarray:
    """Engine<T>::set_R_correlated(groups, rng[, j0, j1])."""
    g = np.ascontiguousarray(groups, dtype=np.int32)
    n = len(g)
    code = CODE_OF[dtype]
    if j0 is None:
        s = int(g.max()) + 1
        cols = s if s < n else n
        a, b = -1, -1
    else:
        cols, a, b = j1 - j0, j0, j1
    R = np.zeros((n, cols), dtype=NP_OF[code], order="F")
    s_out = C.c_int()
    _check(lib().fkpm_host_set_R_correlated(code, n, g.ctypes.data_as(_IP), rng.ptr, a, b,
                                            R.ctypes.data_as(_P), C.byref(s_out)),
           "set_R_correlated")
    return R


def expansion_coefficients(M: int, Mq: int, kind: int, kT: float, mu: float, lo: float,
                           hi: float) -> np.ndarray:
    c = np.zeros(M)
    _check(lib().fkpm_host_expansion_coefficients(M, Mq, kind, kT, mu, lo, hi,
                                                  c.ctypes.data_as(_DP)), "expansion_coefficients")
    return c


def jackson_kernel(M: int) -> np.ndarray:
    g = np.zeros(M)
    lib().fkpm_host_jackson_kernel(M, g.ctypes.data_as(_DP))
    return g


def moment_product(c, mu) -> float:
    c = np.ascontiguousarray(c, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    return float(lib().fkpm_host_moment_product(len(c), c.ctypes.data_as(_DP),
                                                mu.ctypes.data_as(_DP)))


def moment_transform(mu, Mq: int) -> np.ndarray:
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    g = np.zeros(Mq)
    lib().fkpm_host_moment_transform(len(mu), mu.ctypes.data_as(_DP), Mq, g.ctypes.data_as(_DP))
    return g


def electronic_grand_energy(gamma, lo, hi, kT, mu) -> float:
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    return float(lib().fkpm_host_electronic_grand_energy(len(g), g.ctypes.data_as(_DP), lo, hi,
                                                         kT, mu))


def filling_to_mu(gamma, lo, hi, kT, filling, delta_filling=0.0) -> float:
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    return float(lib().fkpm_host_filling_to_mu(len(g), g.ctypes.data_as(_DP), lo, hi, kT,
                                               filling, delta_filling))


def conductivity_coefficients(M: int, Mq: int, kT: float, mu: float, omega: float, lo: float,
                              hi: float, kernel=None) -> np.ndarray:
    """electrical_conductivity_coefficients(M, Mq, kT, mu, omega, es, kernel)
    (Jackson kernel by default): M x M complex, unit e^2/h."""
    k = np.ascontiguousarray(jackson_kernel(M) if kernel is None else kernel, dtype=np.float64)
    out = np.zeros(M * M, dtype=np.complex128)
    _check(lib().fkpm_host_conductivity_coefficients(M, Mq, kT, mu, omega, lo, hi,
                                                     k.ctypes.data_as(_DP),
                                                     out.ctypes.data_as(_DP)),
           "conductivity_coefficients")
    return out.reshape(M, M)


def moment_transform2(mu2, Mq: int, kernel=None) -> np.ndarray:
    """moment_transform of M x M complex moments with a kernel (Jackson by
    default): Mq x Mq complex."""
    m = np.ascontiguousarray(mu2, dtype=np.complex128)
    M = m.shape[0]
    k = np.ascontiguousarray(jackson_kernel(M) if kernel is None else kernel, dtype=np.float64)
    out = np.zeros(Mq * Mq, dtype=np.complex128)
    _check(lib().fkpm_host_moment_transform2(M, m.ctypes.data_as(_DP), Mq, k.ctypes.data_as(_DP),
                                             out.ctypes.data_as(_DP)), "moment_transform2")
    return out.reshape(Mq, Mq)
