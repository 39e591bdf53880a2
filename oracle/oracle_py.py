"""ctypes wrapper of liboracle_fkpm.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker: tests/, the smoke check and the CPU-baseline leg of
bench.py use it to compare against the HIP engine.  Nothing in the product
imports it.  Each function names the oracle.c function (and through it the
reference function) it wraps.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_fkpm.so"
_lib = None

_I, _D, _P, _U32 = C.c_int, C.c_double, C.c_void_p, C.c_uint32
_IP, _DP = C.POINTER(C.c_int), C.POINTER(C.c_double)


def build() -> Path:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        sig = {
            "orc_mt_seed": (None, [_P, _U32]),
            "orc_mt_next": (_U32, [_P]),
            "orc_mt_size": (_I, []),
            "orc_uniform_int": (_I, [_P, _I, _I]),
            "orc_set_R_identity": (None, [_I, _I, _I, _P]),
            "orc_set_R_uncorrelated": (None, [_I, _I, _P, _I, _I, _I, _I, _P]),
            "orc_set_R_correlated": (None, [_I, _IP, _P, _I, _I, _I, _P]),
            "orc_bsr_build": (_I, [_I, _I, _I, _IP, _IP, _P, _IP, _IP, _IP, _P]),
            "orc_bsr_symmetrize": (_I, [_I, _I, _IP, _IP, _IP, _P]),
            "orc_bsr_scale": (None, [_I, _P, _D]),
            "orc_jackson_kernel": (None, [_I, _DP]),
            "orc_chebyshev_fill_array": (None, [_D, _I, _DP, _I]),
            "orc_fermi_energy": (_D, [_D, _D, _D]),
            "orc_fermi_density": (_D, [_D, _D, _D]),
            "orc_chebyshev_points": (None, [_I, _D, _D, _DP]),
            "orc_expansion_coefficients_samples": (None, [_I, _I, _DP, _DP]),
            "orc_expansion_coefficients_fermi": (None, [_I, _I, _I, _D, _D, _D, _D, _DP]),
            "orc_moment_product": (_D, [_I, _DP, _DP]),
            "orc_moment_transform": (None, [_I, _DP, _I, _DP]),
            "orc_electronic_grand_energy": (_D, [_I, _DP, _D, _D, _D, _D]),
            "orc_tridiag_extremes": (None, [_I, _DP, _DP, _DP, _DP]),
            "orc_engine_new": (_P, []),
            "orc_engine_free": (None, [_P]),
            "orc_engine_set_H": (_I, [_P, _I, _I, _I, _IP, _IP, _P, _D, _D]),
            "orc_engine_set_R": (_I, [_P, _I, _I, _P]),
            "orc_engine_moments": (_I, [_P, _I, _DP]),
            "orc_engine_moments_cols": (_I, [_P, _I, _DP, _I, _I]),
            "orc_engine_stoch_matrix": (_I, [_P, _I, _DP, _I, _IP, _IP, _P]),
            "orc_engine_autodiff_matrix": (_I, [_P, _I, _DP, _I, _I, _IP, _IP, _IP, _P]),
            "orc_engine_moments2_v1": (_I, [_P, _I, _I, _IP, _IP, _P, _I, _IP, _IP, _P, _P]),
            "orc_energy_scale": (_I, [_I, _I, _I, _IP, _IP, _P, _D, _I, _P, _DP, _DP, _DP, _DP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _i(a):
    return a.ctypes.data_as(_IP)


def _d(a):
    return a.ctypes.data_as(_DP)


def _p(a):
    return a.ctypes.data_as(_P)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class MT19937:
    """std::mt19937 restated in oracle.c (orc_mt_*)."""

    def __init__(self, seed: int = 5489):
        self._buf = C.create_string_buffer(lib().orc_mt_size())
        lib().orc_mt_seed(self._buf, seed & 0xFFFFFFFF)

    @property
    def ptr(self):
        return C.cast(self._buf, C.c_void_p)

    def __call__(self) -> int:
        return int(lib().orc_mt_next(self.ptr))

    def uniform_int(self, a: int, b: int) -> int:
        return int(lib().orc_uniform_int(self.ptr, a, b))

    def discard(self, k: int) -> None:
        for _ in range(k):
            self()


def set_R_identity(n, j0=0, j1=None):
    j1 = n if j1 is None else j1
    R = np.zeros((j1 - j0) * n, dtype=np.complex128)
    lib().orc_set_R_identity(n, j0, j1, _p(R))
    return R.reshape((j1 - j0, n)).T


def set_R_uncorrelated(n, s, rng: MT19937, j0=0, j1=None, is_complex=True, single=False):
    j1 = s if j1 is None else j1
    R = np.zeros((j1 - j0) * n, dtype=np.complex128)
    lib().orc_set_R_uncorrelated(n, s, rng.ptr, j0, j1, int(is_complex), int(single), _p(R))
    return R.reshape((j1 - j0, n)).T


def set_R_correlated(groups, rng: MT19937, j0=0, j1=None, is_complex=True):
    """Engine::set_R_correlated(groups, rng[, j0, j1]) (engine.cpp:76-116); the
    two-argument form switches to identity when s >= n."""
    g = _i32(groups)
    n = len(g)
    if j1 is None:
        s = int(g.max()) + 1
        if s >= n:
            return set_R_identity(n)
        j0, j1 = 0, s
    R = np.zeros((j1 - j0) * n, dtype=np.complex128)
    lib().orc_set_R_correlated(n, _i(g), rng.ptr, j0, j1, int(is_complex), _p(R))
    return R.reshape((j1 - j0, n)).T


def bsr_build(n_rows, b, rows, cols, vals):
    rows, cols = _i32(rows), _i32(cols)
    v = _c128(vals)
    nb = len(rows)
    rp = np.zeros(n_rows + 1, np.int32)
    ri = np.zeros(max(nb, 1), np.int32)
    ci = np.zeros(max(nb, 1), np.int32)
    val = np.zeros(max(nb, 1) * b * b, np.complex128)
    cnt = lib().orc_bsr_build(n_rows, b, nb, _i(rows), _i(cols), _p(v), _i(rp), _i(ri), _i(ci),
                              _p(val))
    return rp, ri[:cnt], ci[:cnt], val[:cnt * b * b]


def bsr_symmetrize(b, row_ptr, row_idx, col_idx, val):
    val = _c128(val).copy()
    rc = lib().orc_bsr_symmetrize(b, len(row_idx), _i(_i32(row_ptr)), _i(_i32(row_idx)),
                                  _i(_i32(col_idx)), _p(val))
    if rc != 0:
        raise ValueError("pattern is not structurally symmetric")
    return val


def jackson_kernel(M):
    g = np.zeros(M)
    lib().orc_jackson_kernel(M, _d(g))
    return g


def chebyshev_fill_array(x, n, kind=1):
    r = np.zeros(n)
    lib().orc_chebyshev_fill_array(float(x), n, _d(r), kind)
    return r


def fermi_energy(x, kT, mu):
    return lib().orc_fermi_energy(float(x), float(kT), float(mu))


def fermi_density(x, kT, mu):
    return lib().orc_fermi_density(float(x), float(kT), float(mu))


def chebyshev_points(Mq, lo, hi):
    x = np.zeros(Mq)
    lib().orc_chebyshev_points(Mq, lo, hi, _d(x))
    return x


def expansion_coefficients_samples(M, Mq, f_samples):
    f = np.ascontiguousarray(f_samples, dtype=np.float64)
    c = np.zeros(M)
    lib().orc_expansion_coefficients_samples(M, Mq, _d(f), _d(c))
    return c


FERMI_ENERGY, FERMI_DENSITY, CONSTANT_ONE = 0, 1, 2


def expansion_coefficients_fermi(M, Mq, kind, kT, mu, lo, hi):
    c = np.zeros(M)
    lib().orc_expansion_coefficients_fermi(M, Mq, kind, kT, mu, lo, hi, _d(c))
    return c


def moment_product(c, mu):
    c = np.ascontiguousarray(c, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    return lib().orc_moment_product(len(c), _d(c), _d(mu))


def moment_transform(mu, Mq):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    g = np.zeros(Mq)
    lib().orc_moment_transform(len(mu), _d(mu), Mq, _d(g))
    return g


def electronic_grand_energy(gamma, lo, hi, kT, mu):
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    return lib().orc_electronic_grand_energy(len(g), _d(g), lo, hi, kT, mu)


def tridiag_extremes(alpha, beta):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    lo, hi = C.c_double(), C.c_double()
    lib().orc_tridiag_extremes(len(a), _d(a), _d(b), C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def energy_scale(H, extend, iters, v_start):
    v = _c128(v_start)
    al = np.zeros(iters)
    be = np.zeros(iters)
    lo, hi = C.c_double(), C.c_double()
    lib().orc_energy_scale(H.n_rows, H.b_len, H.n_blocks(), _i(_i32(H.row_ptr)),
                           _i(_i32(H.col_idx)), _p(_c128(H.val)), float(extend), int(iters),
                           _p(v), _d(al), _d(be), C.byref(lo), C.byref(hi))
    return lo.value, hi.value, al, be


class Engine:
    """Engine_CPU (engine_cpu.cpp) restated in oracle.c, FP64 complex."""

    def __init__(self):
        self._e = lib().orc_engine_new()
        self.b = 0

    def __del__(self):
        try:
            lib().orc_engine_free(self._e)
        except Exception:
            pass

    def set_H(self, H, lo, hi):
        self.b = H.b_len
        self._H = H
        rc = lib().orc_engine_set_H(self._e, H.n_rows, H.b_len, H.n_blocks(), _i(_i32(H.row_ptr)),
                                    _i(_i32(H.col_idx)), _p(_c128(H.val)), float(lo), float(hi))
        if rc != 0:
            raise ValueError("oracle set_H: missing diagonal block")

    def set_R(self, R):
        R = np.asfortranarray(R, dtype=np.complex128)
        n, s = R.shape
        lib().orc_engine_set_R(self._e, n, s, R.ctypes.data_as(_P))

    def moments(self, M):
        mu = np.zeros(M)
        if lib().orc_engine_moments(self._e, M, _d(mu)) != 0:
            raise ValueError("oracle moments failed")
        return mu

    def moments_cols(self, M, c0, c1):
        mu = np.zeros(M)
        if lib().orc_engine_moments_cols(self._e, M, _d(mu), c0, c1) != 0:
            raise ValueError("oracle moments_cols failed")
        return mu

    def stoch_matrix(self, c, D):
        c = np.ascontiguousarray(c, dtype=np.float64)
        out = np.zeros(D.n_blocks() * D.b_len ** 2, np.complex128)
        lib().orc_engine_stoch_matrix(self._e, len(c), _d(c), D.n_blocks(), _i(_i32(D.row_idx)),
                                      _i(_i32(D.col_idx)), _p(out))
        return out

    def moments2_v1(self, M, j1, j2):
        """Engine_CPU::moments2_v1 (engine_cpu.cpp:86-160): M x M complex."""
        mu = np.zeros(M * M, np.complex128)
        v1 = np.ascontiguousarray(j1.val, dtype=np.complex128)
        v2 = np.ascontiguousarray(j2.val, dtype=np.complex128)
        rc = lib().orc_engine_moments2_v1(self._e, M, j1.n_blocks(), _i(_i32(j1.row_ptr)),
                                          _i(_i32(j1.col_idx)), _p(v1), j2.n_blocks(),
                                          _i(_i32(j2.row_ptr)), _i(_i32(j2.col_idx)), _p(v2),
                                          _p(mu))
        if rc != 0:
            raise ValueError(f"oracle moments2_v1 failed ({rc})")
        return mu.reshape(M, M)

    def autodiff_matrix(self, c, D):
        c = np.ascontiguousarray(c, dtype=np.float64)
        out = np.zeros(D.n_blocks() * D.b_len ** 2, np.complex128)
        rc = lib().orc_engine_autodiff_matrix(self._e, len(c), _d(c), D.n_rows, D.n_blocks(),
                                              _i(_i32(D.row_ptr)), _i(_i32(D.row_idx)),
                                              _i(_i32(D.col_idx)), _p(out))
        if rc != 0:
            raise ValueError(f"oracle autodiff failed ({rc})")
        return out
