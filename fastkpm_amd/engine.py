"""Python mirror of the fkpm::Engine<T> interface (src/fastkpm.h:185-234) over
the C-ABI of libfastkpm_hip.so.

Same method names, argument meaning and call-order rules as the reference
engine: set_R_* / set_R (transfer_R), set_H(H, es), moments(M),
stoch_matrix(c, D), autodiff_matrix(c, D) (which requires moments(M) with the
same M just before).  Errors raise FkpmError instead of aborting.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import call, lib

_DTYPES = {
    "float": (_capi.F32, np.float32, False),
    "double": (_capi.F64, np.float64, False),
    "cx_float": (_capi.C32, np.complex64, True),
    "cx_double": (_capi.C64, np.complex128, True),
}
_PRECS = {"fp64": _capi.PREC_FP64, "fp32": _capi.PREC_FP32}


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class EnergyScale:
    """fkpm::EnergyScale (src/fastkpm.h:112-122)."""
    lo: float
    hi: float

    def avg(self) -> float:
        return (self.hi + self.lo) / 2.0

    def mag(self) -> float:
        return (self.hi - self.lo) / 2.0

    def scale(self, x):
        return (x - self.avg()) / self.mag()

    def unscale(self, x):
        return x * self.mag() + self.avg()


@dataclass
class SpMatBsr:
    """Block sparse row matrix, the fields of fkpm::SpMatBsr (src/fastkpm.h:66-74).

    val holds nnzb*b*b values, each b x b block column-major (element (bi, bj)
    of block k at k*b*b + bj*b + bi), block rows sorted by column.
    """
    n_rows: int
    n_cols: int
    b_len: int
    row_idx: np.ndarray
    col_idx: np.ndarray
    row_ptr: np.ndarray
    val: np.ndarray = field(default=None)

    def n_blocks(self) -> int:
        return int(len(self.row_idx))

    def zeros(self) -> None:
        self.val[:] = 0

    def like(self, dtype=None) -> "SpMatBsr":
        """Same pattern, zero values (the caller-owned D of stoch/autodiff)."""
        dt = self.val.dtype if dtype is None else dtype
        return SpMatBsr(self.n_rows, self.n_cols, self.b_len, self.row_idx.copy(),
                        self.col_idx.copy(), self.row_ptr.copy(),
                        np.zeros(self.n_blocks() * self.b_len ** 2, dtype=dt))

    def find_index(self, i: int, j: int) -> int:
        lo, hi = int(self.row_ptr[i]), int(self.row_ptr[i + 1])
        hits = np.nonzero(self.col_idx[lo:hi] == j)[0]
        if len(hits) == 0:
            raise KeyError(f"could not find block index ({i},{j}) in BSR matrix")
        return lo + int(hits[0])

    def block(self, k: int) -> np.ndarray:
        b = self.b_len
        return self.val[k * b * b:(k + 1) * b * b].reshape(b, b).T

    def to_dense(self) -> np.ndarray:
        b = self.b_len
        out = np.zeros((self.n_rows * b, self.n_cols * b), dtype=self.val.dtype)
        for k in range(self.n_blocks()):
            i, j = int(self.row_idx[k]), int(self.col_idx[k])
            out[b * i:b * i + b, b * j:b * j + b] += self.block(k)
        return out


class HipEngine:
    """fkpm::Engine<T> on one HIP device.

    dtype: 'float', 'double', 'cx_float', 'cx_double' (the engine's T).
    prec:  'fp64' (vectors and H in FP64) or 'fp32' (FP32 recursion storage
           with FP64 row sums, moment dots, xi and D accumulation).
    """

    def __init__(self, dtype: str = "cx_double", prec: str = "fp64", device: int = 0):
        if dtype not in _DTYPES:
            raise ValueError(f"dtype must be one of {list(_DTYPES)}")
        self.dtype = dtype
        self._code, self.np_dtype, self.is_complex = _DTYPES[dtype]
        self.prec = prec
        h = C.c_void_p()
        call("fkpm_dev_open", device, self._code, _PRECS[prec], C.byref(h))
        self._h = h
        self.R = None
        self.es = None
        self._d_pattern_key = None

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fkpm_dev_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def stream(self) -> int:
        return int(lib().fkpm_stream(self._h) or 0)

    def synchronize(self) -> None:
        call("fkpm_synchronize", self._h)

    # -- R ------------------------------------------------------------------
    def set_R(self, R: np.ndarray) -> None:
        """Set Engine::R (n x s) and transfer it to the device (transfer_R)."""
        R = np.asfortranarray(R, dtype=self.np_dtype)
        n, s = R.shape
        self.R = R
        call("fkpm_set_R", self._h, n, s, R.ctypes.data_as(C.c_void_p))

    def set_R_identity(self, n: int, j_start: int = 0, j_end: int | None = None) -> None:
        from . import host
        self.set_R(host.set_R_identity(self.dtype, n, j_start, j_end))

    def set_R_uncorrelated(self, n: int, s: int, rng, j_start: int = 0,
                           j_end: int | None = None) -> None:
        from . import host
        self.set_R(host.set_R_uncorrelated(self.dtype, n, s, rng, j_start, j_end))

    def set_R_correlated(self, groups, rng, j_start: int | None = None,
                         j_end: int | None = None) -> None:
        """Engine::set_R_correlated (src/engine.cpp:76-116), drawing from the
        C++ library's std::mt19937 `rng` (fastkpm_amd.host.RNG)."""
        from . import host
        self.set_R(host.set_R_correlated(self.dtype, groups, rng, j_start, j_end))

    # -- spectrum -----------------------------------------------------------
    def energy_scale(self, H: SpMatBsr, extend: float, iters: int = 128,
                     v_start: np.ndarray | None = None) -> EnergyScale:
        """Lanczos bounds of H (Engine::energy_scale, src/fastkpm.h:193-194)."""
        n = H.n_rows * H.b_len
        if v_start is None:
            rs = np.random.default_rng(0)
            v_start = rs.standard_normal(n)
            if self.is_complex:
                v_start = v_start + 1j * rs.standard_normal(n)
        v = np.ascontiguousarray(v_start, dtype=self.np_dtype)
        val = np.ascontiguousarray(H.val, dtype=self.np_dtype)
        rp = np.ascontiguousarray(H.row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(H.col_idx, dtype=np.int32)
        lo, hi = C.c_double(), C.c_double()
        call("fkpm_energy_scale", self._h, H.n_rows, H.b_len, H.n_blocks(), _iptr(rp), _iptr(ci),
             val.ctypes.data_as(C.c_void_p), float(extend), int(iters),
             v.ctypes.data_as(C.c_void_p), C.byref(lo), C.byref(hi))
        return EnergyScale(lo.value, hi.value)

    def set_H(self, H: SpMatBsr, es: EnergyScale) -> None:
        if H.n_rows != H.n_cols:
            raise ValueError("H must be square")
        self.es = es
        val = np.ascontiguousarray(H.val, dtype=self.np_dtype)
        rp = np.ascontiguousarray(H.row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(H.col_idx, dtype=np.int32)
        call("fkpm_set_H", self._h, H.n_rows, H.b_len, H.n_blocks(), _iptr(rp), _iptr(ci),
             val.ctypes.data_as(C.c_void_p), float(es.avg()), float(es.mag()))

    # -- recursions ---------------------------------------------------------
    def moments(self, M: int) -> np.ndarray:
        mu = np.zeros(M, dtype=np.float64)
        call("fkpm_moments", self._h, int(M), _dptr(mu))
        return mu

    def _set_D(self, D: SpMatBsr) -> None:
        rp = np.ascontiguousarray(D.row_ptr, dtype=np.int32)
        ri = np.ascontiguousarray(D.row_idx, dtype=np.int32)
        ci = np.ascontiguousarray(D.col_idx, dtype=np.int32)
        key = (D.n_rows, rp.tobytes(), ci.tobytes())
        if key != self._d_pattern_key:
            call("fkpm_set_D_pattern", self._h, D.n_rows, D.n_blocks(), _iptr(rp), _iptr(ri),
                 _iptr(ci))
            self._d_pattern_key = key

    def _D_out(self, D: SpMatBsr, keep_on_device: bool):
        if keep_on_device:
            return None
        if D.val is None or D.val.dtype != self.np_dtype or not D.val.flags.c_contiguous \
                or D.val.size != D.n_blocks() * D.b_len ** 2:
            D.val = np.zeros(D.n_blocks() * D.b_len ** 2, dtype=self.np_dtype)
        return D.val.ctypes.data_as(C.c_void_p)

    def stoch_matrix(self, c, D: SpMatBsr, keep_on_device: bool = False) -> None:
        c = np.ascontiguousarray(c, dtype=np.float64)
        self._set_D(D)
        call("fkpm_stoch_matrix", self._h, len(c), _dptr(c), self._D_out(D, keep_on_device))

    def autodiff_matrix(self, c, D: SpMatBsr, keep_on_device: bool = False) -> None:
        c = np.ascontiguousarray(c, dtype=np.float64)
        self._set_D(D)
        call("fkpm_autodiff_matrix", self._h, len(c), _dptr(c), self._D_out(D, keep_on_device))

    def get_D(self, D: SpMatBsr) -> None:
        call("fkpm_get_D", self._h, self._D_out(D, False))

    # -- introspection ------------------------------------------------------
    def set_timing(self, on: bool) -> None:
        call("fkpm_set_timing", self._h, 1 if on else 0)

    def timing(self) -> dict:
        v = [C.c_double(), C.c_int(), C.c_double(), C.c_int(), C.c_double(), C.c_int()]
        call("fkpm_get_timing", self._h, *[C.byref(x) for x in v])
        return {"moments_ms": v[0].value, "moments_launches": v[1].value,
                "stoch_ms": v[2].value, "stoch_launches": v[3].value,
                "autodiff_ms": v[4].value, "autodiff_launches": v[5].value}

    def config(self) -> dict:
        v = [C.c_int() for _ in range(5)]
        call("fkpm_get_config", self._h, *[C.byref(x) for x in v])
        return {"lanes": v[0].value, "kq": v[1].value, "s_padded": v[2].value,
                "n_wg": v[3].value, "fused_autodiff": bool(v[4].value)}

    def set_geometry(self, lanes: int = 0, kq: int = 0, max_wg: int = 0) -> None:
        call("fkpm_set_geometry", self._h, int(lanes), int(kq), int(max_wg))

    def autotune(self) -> None:
        """Time the kernel geometries on the current H and R and keep the fastest."""
        call("fkpm_autotune", self._h)

    def set_stoch_method(self, method: str) -> None:
        """'auto' (default: from the forward vectors a preceding moments kept,
        else Clenshaw), 'clenshaw' or 'forward' (the reference's summation
        order)."""
        v = {"auto": _capi.STOCH_AUTO, "clenshaw": _capi.STOCH_CLENSHAW,
             "forward": _capi.STOCH_FORWARD}[method]
        call("fkpm_set_option", self._h, _capi.OPT_STOCH_METHOD, v)

    def stoch_method(self) -> str:
        """The method the last stoch_matrix ran: 'clenshaw', 'forward',
        'table' or '' (none yet)."""
        v = C.c_int()
        call("fkpm_get_stoch_method", self._h, C.byref(v))
        return {_capi.STOCH_CLENSHAW: "clenshaw", _capi.STOCH_FORWARD: "forward",
                _capi.STOCH_TABLE: "table"}.get(v.value, "")

    def set_autotune_on_use(self, on: bool) -> None:
        call("fkpm_set_option", self._h, _capi.OPT_AUTOTUNE, 1 if on else 0)

    def moments2_v1(self, M: int, j1: "SpMatBsr", j2: "SpMatBsr", col_chunk: int = 0) -> np.ndarray:
        """Two-current moments mu[m1][m2] = sum_k <T_m1(Hs) r_k, j2 T_m2(Hs) j1 r_k>
        (Engine::moments2_v1): M x M complex.  j1, j2 on H's block rows, the
        engine's element type; col_chunk bounds the probe columns per pass."""
        out = np.zeros(M * M * 2)
        args = []
        keep = []
        for j in (j1, j2):
            rp = np.ascontiguousarray(j.row_ptr, dtype=np.int32)
            ci = np.ascontiguousarray(j.col_idx, dtype=np.int32)
            v = np.ascontiguousarray(j.val, dtype=self.np_dtype)
            keep += [rp, ci, v]
            args += [j.n_blocks(), _iptr(rp), _iptr(ci), v.ctypes.data_as(C.c_void_p)]
        call("fkpm_moments2", self._h, int(M), *args, int(col_chunk), _dptr(out))
        return out.view(np.complex128).reshape(M, M)

    def set_store_alpha(self, mode) -> None:
        """Keep every forward vector for autodiff_matrix (FKPM_OPT_STORE_ALPHA):
        True / 2 whenever the table fits, 1 automatic (default), False / 0
        recompute alpha backward."""
        call("fkpm_set_option", self._h, _capi.OPT_STORE_ALPHA,
             2 if mode is True else 0 if mode is False else int(mode))

    def alpha_table(self) -> bool:
        """Whether the last moments(M) kept its forward vectors."""
        v = C.c_int()
        call("fkpm_get_alpha_table", self._h, C.byref(v))
        return bool(v.value)

    def set_sweep(self, sweep: int) -> None:
        """Row sweep of every kernel kind: 0 per XCD, 1 one front, -1 the
        configuration's own (FKPM_OPT_SWEEP)."""
        call("fkpm_set_option", self._h, _capi.OPT_SWEEP, int(sweep))

    def set_slice(self, mode: int) -> None:
        """Column-pass loop order of the vector-unit step kernels
        (FKPM_OPT_SLICE): 1 every row once per column pass, 0 a row's passes
        back to back, -1 the configuration's own."""
        call("fkpm_set_option", self._h, _capi.OPT_SLICE, int(mode))

    def set_row_order(self, mode: int) -> None:
        """Lattice row order of the vector-unit step kernels
        (FKPM_OPT_ROW_ORDER): 1 slab by slab in y-tiles along z when H is a
        periodic 3-D lattice, 0 row order, -1 the configuration's own."""
        call("fkpm_set_option", self._h, _capi.OPT_ROW_ORDER, int(mode))

    def set_graph(self, mode: int) -> None:
        """Step loops as hipGraphs (FKPM_OPT_GRAPH): 1 always, 0 never, -1
        automatic (short steps)."""
        call("fkpm_set_option", self._h, _capi.OPT_GRAPH, int(mode))

    def graph_stats(self) -> dict:
        a, b = C.c_longlong(), C.c_longlong()
        call("fkpm_get_graph_stats", self._h, C.byref(a), C.byref(b))
        return {"graph_launches": a.value, "graph_captures": b.value}

    def set_alpha_table_mb(self, mib: int) -> None:
        """Most MiB the table of forward vectors may take on this handle
        (FKPM_OPT_ALPHA_TABLE_MB; -1: a quarter of the device, at most half
        of what is free)."""
        call("fkpm_set_option", self._h, _capi.OPT_ALPHA_TABLE_MB, int(mib))

    def set_mfma(self, on: bool | int) -> None:
        """Matrix-core kernels where the block shape has them (complex 8x8
        blocks; on by default), or the vector-unit kernels only (False); 2:
        matrix cores with the row-resident kernels as the untuned choice of
        every step that has one (FKPM_OPT_MFMA)."""
        mode = on if isinstance(on, int) and not isinstance(on, bool) else (1 if on else 0)
        call("fkpm_set_option", self._h, _capi.OPT_MFMA, int(mode))

    def kernel_config(self, kind: int) -> dict:
        v = [C.c_int() for _ in range(4)]
        call("fkpm_get_kernel_config", self._h, int(kind), *[C.byref(x) for x in v])
        mf, sw = C.c_int(), C.c_int()
        call("fkpm_get_kernel_mfma", self._h, int(kind), C.byref(mf))
        call("fkpm_get_kernel_sweep", self._h, int(kind), C.byref(sw))
        return {"lanes": v[0].value, "kq": v[1].value, "n_wg": v[2].value, "nt": v[3].value,
                "mfma": bool(mf.value), "mfma_variant": mf.value, "sweep": sw.value}

    def state(self) -> dict:
        v = [C.c_int() for _ in range(3)]
        call("fkpm_get_state", self._h, *[C.byref(x) for x in v])
        return {"r_compact": bool(v[0].value), "tuned": bool(v[1].value),
                "moments_M": v[2].value}


def bandwidth_probe(eng: "HipEngine", nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Read+write GB/s of the faster of two copy kernels (plain, and four
    packets per thread with nontemporal hints) over nbytes on eng's device."""
    g = C.c_double()
    call("fkpm_bandwidth_probe", eng.handle, int(nbytes), int(reps), C.byref(g))
    return g.value


def lattice_order(H: SpMatBsr):
    """The lattice row order set_H builds for H's pattern (FKPM_OPT_ROW_ORDER):
    ((n1, n2, n3), order) when H is a periodic 3-D lattice in natural order,
    else None.  A host computation; no device needed."""
    nb = int(H.n_rows)
    rp = np.ascontiguousarray(H.row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(H.col_idx, dtype=np.int32)
    order = np.empty(nb, dtype=np.int32)
    dims = np.zeros(3, dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    rc = lib().fkpm_lattice_order(nb, int(len(ci)), ip(rp), ip(ci), ip(order), ip(dims))
    if rc < 0:
        _capi.check(rc, "fkpm_lattice_order")
    return (tuple(int(d) for d in dims), order) if rc == 1 else None


def device_count() -> int:
    return int(lib().fkpm_device_count())
