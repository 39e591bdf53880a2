"""ctypes binding of the C-ABI in include/fastkpm_hip.h (libfastkpm_hip.so).

Loading fails loudly when the library has not been built: there is no CPU
fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libfastkpm_hip.so"

F32, F64, C32, C64 = 0, 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1
OPT_STOCH_METHOD, OPT_AUTOTUNE, OPT_MFMA, OPT_SWEEP, OPT_STORE_ALPHA = 1, 2, 3, 4, 5
OPT_SLICE, OPT_ALPHA_TABLE_MB, OPT_GRAPH, OPT_ROW_ORDER = 6, 7, 8, 9
STOCH_CLENSHAW, STOCH_FORWARD, STOCH_TABLE, STOCH_AUTO = 0, 1, 2, 3
KERNEL_MOMENTS, KERNEL_XI, KERNEL_AUTODIFF, KERNEL_OUTER = 0, 1, 2, 3
KERNEL_MOMENTS_TABLE, KERNEL_AUTODIFF_TABLE = 4, 5

ERRORS = {-1: "FKPM_E_ARG", -2: "FKPM_E_STATE", -3: "FKPM_E_HIP", -4: "FKPM_E_NODEV",
          -5: "FKPM_E_UNSUP"}

# Every entry point of include/fastkpm_hip.h with its ctypes signature.
_I, _D, _P = C.c_int, C.c_double, C.c_void_p
_IP, _DP = C.POINTER(C.c_int), C.POINTER(C.c_double)
SIGNATURES = {
    "fkpm_last_error": (C.c_char_p, []),
    "fkpm_device_count": (_I, []),
    "fkpm_dev_open": (_I, [_I, _I, _I, C.POINTER(_P)]),
    "fkpm_dev_close": (_I, [_P]),
    "fkpm_stream": (_P, [_P]),
    "fkpm_synchronize": (_I, [_P]),
    "fkpm_energy_scale": (_I, [_P, _I, _I, _I, _IP, _IP, _P, _D, _I, _P, _DP, _DP]),
    "fkpm_set_H": (_I, [_P, _I, _I, _I, _IP, _IP, _P, _D, _D]),
    "fkpm_set_R": (_I, [_P, _I, _I, _P]),
    "fkpm_moments": (_I, [_P, _I, _DP]),
    "fkpm_set_D_pattern": (_I, [_P, _I, _I, _IP, _IP, _IP]),
    "fkpm_stoch_matrix": (_I, [_P, _I, _DP, _P]),
    "fkpm_autodiff_matrix": (_I, [_P, _I, _DP, _P]),
    "fkpm_get_D": (_I, [_P, _P]),
    "fkpm_moments2": (_I, [_P, _I, _I, _IP, _IP, _P, _I, _IP, _IP, _P, _I, _DP]),
    "fkpm_set_timing": (_I, [_P, _I]),
    "fkpm_get_timing": (_I, [_P, _DP, _IP, _DP, _IP, _DP, _IP]),
    "fkpm_get_config": (_I, [_P, _IP, _IP, _IP, _IP, _IP]),
    "fkpm_set_geometry": (_I, [_P, _I, _I, _I]),
    "fkpm_bandwidth_probe": (_I, [_P, C.c_size_t, _I, _DP]),
    "fkpm_set_option": (_I, [_P, _I, _I]),
    "fkpm_autotune": (_I, [_P]),
    "fkpm_get_kernel_config": (_I, [_P, _I, _IP, _IP, _IP, _IP]),
    "fkpm_get_kernel_mfma": (_I, [_P, _I, _IP]),
    "fkpm_get_kernel_sweep": (_I, [_P, _I, _IP]),
    "fkpm_get_alpha_table": (_I, [_P, _IP]),
    "fkpm_get_stoch_method": (_I, [_P, _IP]),
    "fkpm_get_state": (_I, [_P, _IP, _IP, _IP]),
    "fkpm_get_graph_stats": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "fkpm_lattice_order": (_I, [_I, _I, _IP, _IP, _IP, _IP]),
    "fkpm_moments_noalloc": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _D, _D, _I, _I, _P, _I, _P,
                                  _P]),
    # batched engine
    "fkpm_batch_footprint": (_I, [_I, _I, _I, _I, _I, _I, C.POINTER(C.c_size_t)]),
    "fkpm_batch_open": (_I, [_I, _I, C.c_size_t, _I, C.POINTER(_P)]),
    "fkpm_batch_close": (_I, [_P]),
    "fkpm_batch_set_option": (_I, [_P, _I, _I]),
    "fkpm_batch_submit": (_I, [_P, _I, _I, _I, _IP, _IP, _P, _D, _D, _I, _I, _P, _I, _DP,
                               C.POINTER(C.c_longlong)]),
    "fkpm_batch_submit_probing": (_I, [_P, _I, _I, _I, _IP, _IP, _P, _D, _D, _I, _I, _IP, _P,
                                       _I, _DP, C.POINTER(C.c_longlong)]),
    "fkpm_batch_submit_seeded": (_I, [_P, _I, _I, _I, _IP, _IP, _P, _D, _D, _I, _I, _IP,
                                      C.POINTER(C.c_uint), _I, _DP, C.POINTER(C.c_longlong)]),
    "fkpm_batch_poll": (_I, [_P, C.c_longlong, _IP]),
    "fkpm_batch_wait": (_I, [_P, C.c_longlong, _DP, _DP, _DP, _DP]),
    "fkpm_batch_stats": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                              C.POINTER(C.c_longlong), _IP, C.POINTER(C.c_size_t)]),
}
BATCH_OPT_GRAPH, BATCH_OPT_TICKS = 1, 2

_lib = None


class FkpmError(RuntimeError):
    """A C-ABI call returned a nonzero status."""

    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} -> {ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """The loaded device library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                "g.build()' or python fastkpm_amd/build.py). There is no CPU fallback.")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, fn: str) -> None:
    if rc != 0:
        msg = lib().fkpm_last_error().decode(errors="replace")
        raise FkpmError(rc, fn, msg)


def call(fn: str, *args) -> None:
    check(getattr(lib(), fn)(*args), fn)
