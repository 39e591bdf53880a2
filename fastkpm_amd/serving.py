"""Batched engine (serving runtime) over the C-ABI of libfastkpm_hip.so, and
the mixed request stream it is measured on.

A request is one energy-and-force solve of a simulation loop: moments(M),
E = sum_m c_m mu_m with c from expansion_coefficients(fermi_energy), and
dE/dH = autodiff_matrix(c, D) on H's own pattern -- the sequence of
Engine::moments / autodiff_matrix (src/engine_cpu.cpp:63-84, 255-304) for one
Hamiltonian.  BatchEngine queues requests; the library's scheduler thread
admits them by HBM footprint and runs all admitted ones together.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _capi, host, lattice
from ._capi import call
from .engine import EnergyScale, SpMatBsr

_PRECS = {"fp64": _capi.PREC_FP64, "fp32": _capi.PREC_FP32}


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def footprint(prec: str, H: SpMatBsr, S: int, M: int) -> int:
    """HBM bytes a request holds while admitted."""
    out = C.c_size_t()
    call("fkpm_batch_footprint", _PRECS[prec], H.n_rows, H.b_len, H.n_blocks(), int(S), int(M),
         C.byref(out))
    return out.value


@dataclass
class Result:
    mu: np.ndarray        # moments mu_0 .. mu_{M-1}
    energy: float         # sum_m c_m mu_m
    D: np.ndarray | None  # dE/dH on H's pattern (nnzb, b, b) complex, block column-major values
    queued_ms: float
    run_ms: float
    total_ms: float


class BatchEngine:
    """Handle of one batched engine (fkpm_batch_* of fastkpm_hip.h)."""

    def __init__(self, prec: str = "fp64", hbm_bytes: int = 16 << 30, max_slots: int = 64,
                 device: int = 0, graph: bool = True, max_ticks: int = 16):
        self.prec = prec
        h = C.c_void_p()
        call("fkpm_batch_open", int(device), _PRECS[prec], int(hbm_bytes), int(max_slots),
             C.byref(h))
        self._h = h
        self._meta = {}
        call("fkpm_batch_set_option", self._h, _capi.BATCH_OPT_GRAPH, 1 if graph else 0)
        call("fkpm_batch_set_option", self._h, _capi.BATCH_OPT_TICKS, int(max_ticks))

    def close(self) -> None:
        if self._h:
            call("fkpm_batch_close", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, H: SpMatBsr, es: EnergyScale, c, R=None, probing=None,
               seeded=None) -> int:
        """Queue a request; R dense (n x s), probing = (rcol, rval, s) or
        seeded = (groups, seeds): colours per row and one generator seed per
        column (master_seeds), the phases then drawn on the device."""
        rp = np.ascontiguousarray(H.row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(H.col_idx, dtype=np.int32)
        val = np.ascontiguousarray(H.val, dtype=np.complex128)
        c = np.ascontiguousarray(c, dtype=np.float64)
        n = H.n_rows * H.b_len
        rid = C.c_longlong()
        if seeded is not None:
            groups, seeds = seeded
            groups = np.ascontiguousarray(groups, dtype=np.int32)
            seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
            call("fkpm_batch_submit_seeded", self._h, H.n_rows, H.b_len, H.n_blocks(), _ip(rp),
                 _ip(ci), val.ctypes.data_as(C.c_void_p), float(es.lo), float(es.hi), n,
                 len(seeds), _ip(groups), seeds.ctypes.data_as(C.POINTER(C.c_uint)), len(c),
                 _dp(c), C.byref(rid))
        elif probing is not None:
            rcol, rval, s = probing
            rcol = np.ascontiguousarray(rcol, dtype=np.int32)
            rval = np.ascontiguousarray(rval, dtype=np.complex128)
            call("fkpm_batch_submit_probing", self._h, H.n_rows, H.b_len, H.n_blocks(), _ip(rp),
                 _ip(ci), val.ctypes.data_as(C.c_void_p), float(es.lo), float(es.hi), n, int(s),
                 _ip(rcol), rval.ctypes.data_as(C.c_void_p), len(c), _dp(c), C.byref(rid))
        else:
            R = np.asfortranarray(R, dtype=np.complex128)
            s = R.shape[1]
            call("fkpm_batch_submit", self._h, H.n_rows, H.b_len, H.n_blocks(), _ip(rp), _ip(ci),
                 val.ctypes.data_as(C.c_void_p), float(es.lo), float(es.hi), n, int(s),
                 R.ctypes.data_as(C.c_void_p), len(c), _dp(c), C.byref(rid))
        self._meta[rid.value] = (len(c), H.n_blocks(), H.b_len)
        return rid.value

    def poll(self, rid: int) -> bool:
        d = C.c_int()
        call("fkpm_batch_poll", self._h, C.c_longlong(rid), C.byref(d))
        return bool(d.value)

    def wait(self, rid: int, want_D: bool = True) -> Result:
        M, nnzb, b = self._meta.pop(rid)
        mu = np.zeros(M)
        e = C.c_double()
        t = np.zeros(3)
        D = np.zeros(nnzb * b * b, dtype=np.complex128) if want_D else None
        call("fkpm_batch_wait", self._h, C.c_longlong(rid), _dp(mu), C.byref(e),
             D.ctypes.data_as(C.POINTER(C.c_double)) if want_D else None, _dp(t))
        return Result(mu, e.value, D, t[0], t[1], t[2])

    def stats(self) -> dict:
        v = [C.c_longlong() for _ in range(3)]
        ps, pb = C.c_int(), C.c_size_t()
        call("fkpm_batch_stats", self._h, *[C.byref(x) for x in v], C.byref(ps), C.byref(pb))
        return {"ticks": v[0].value, "replays": v[1].value, "captures": v[2].value,
                "peak_slots": ps.value, "peak_bytes": pb.value}


# ---------------------------------------------------------------------------
# The mixed request stream: 2-D / 3-D lattices, 2-spin (Anderson + Kondo) or
# 8-orbital blocks, N from 4096 to 1048576, M from 256 to 2048, S from 16 to
# 128 probing colours, disorder W uniform in [0, 6], seeded.
# ---------------------------------------------------------------------------

def gershgorin(H: SpMatBsr, pad: float = 0.01) -> EnergyScale:
    """Spectral bounds of a Hermitian BSR matrix from Gershgorin discs."""
    b = H.b_len
    blocks = H.val.reshape(-1, b, b, order="C")  # [p][bj][bi] (column-major blocks)
    absrow = np.abs(blocks).sum(axis=1)          # sum over bj -> [p][bi]
    rows = np.repeat(np.arange(H.n_rows), np.diff(H.row_ptr))
    radius = np.zeros((H.n_rows, b))
    np.add.at(radius, rows, absrow)
    diag_p = np.flatnonzero(rows == np.asarray(H.col_idx))
    d = np.zeros((H.n_rows, b))
    d[rows[diag_p]] = np.real(np.diagonal(blocks[diag_p], axis1=1, axis2=2))
    radius[rows[diag_p]] -= np.abs(np.diagonal(blocks[diag_p], axis1=1, axis2=2))
    lo = float(np.min(d - radius))
    hi = float(np.max(d + radius))
    w = hi - lo
    return EnergyScale(lo - pad * w, hi + pad * w)


_SHAPES = {
    # (dims, b) -> (edge lengths giving N in [4096, 1048576], colour period choices)
    (2, 2): (range(46, 725), (3, 4, 5, 6, 7, 8)),
    (3, 2): (range(13, 81), (2, 3, 4)),
    (2, 8): (range(23, 363), (2, 3, 4)),
    (3, 8): (range(8, 51), (2,)),
}


@dataclass
class RequestSpec:
    dims: int
    b: int
    edge: int
    period: int
    M: int
    W: float
    seed: int

    @property
    def shape(self):
        return (self.edge,) * self.dims

    @property
    def N(self) -> int:
        return self.edge ** self.dims * self.b

    @property
    def S(self) -> int:
        return self.b * self.period ** self.dims


def draw_spec(rng: np.random.Generator, n_max: int = 1 << 20, m_max: int = 2048) -> RequestSpec:
    """One request of the mixed stream (log-uniform N and M)."""
    dims, b = [(2, 2), (3, 2), (2, 8), (3, 8)][rng.integers(4)]
    edges, periods = _SHAPES[(dims, b)]
    logn = rng.uniform(math.log(4096), math.log(min(n_max, 1 << 20)))
    edge = int(round((math.exp(logn) / b) ** (1.0 / dims)))
    edge = min(max(edge, edges.start), edges.stop - 1)
    while edge ** dims * b > n_max and edge > 2:
        edge -= 1
    period = int(rng.choice([p for p in periods if edge % p == 0] or [periods[0]]))
    M = int(round(math.exp(rng.uniform(math.log(256), math.log(max(256, m_max)))) / 8)) * 8
    W = float(rng.uniform(0.0, 6.0))
    return RequestSpec(dims, b, edge, period, M, W, int(rng.integers(1 << 30)))


def build_request(spec: RequestSpec, kT: float = 0.05, mu: float = 0.0):
    """H, energy scale, coefficients and the probing vectors of a spec as
    colours plus column seeds (drawn on the device at admission)."""
    if spec.b == 2:
        H = lattice.lattice(spec.shape, b=2, t=-1.0, J=1.0, W=spec.W, seed=spec.seed)
    else:
        H = lattice.lattice(spec.shape, b=8, orbital_blocks=True, seed=spec.seed)
        if spec.W:
            # on-site disorder eps_i * I on the orbital blocks
            eps = np.random.default_rng(spec.seed + 1).uniform(-spec.W / 2, spec.W / 2, H.n_rows)
            diag_p = np.flatnonzero(np.repeat(np.arange(H.n_rows), np.diff(H.row_ptr))
                                    == np.asarray(H.col_idx))
            blk = H.val.reshape(-1, 8, 8)
            blk[diag_p] += eps[:, None, None] * np.eye(8)[None]
    es = gershgorin(H)
    c = host.expansion_coefficients(spec.M, 2 * spec.M, host.FERMI_ENERGY, kT, mu, es.lo, es.hi)
    groups = lattice.colouring(spec.shape, (spec.period,) * spec.dims, spec.b)
    seeded = (groups, host.master_seeds(host.RNG(spec.seed), int(groups.max()) + 1))
    return H, es, c, seeded


def run_stream(eng: BatchEngine, specs, rate_hz: float, seed: int = 0, want_D: bool = False,
               log=None, built=None):
    """Submit the requests at Poisson arrival times (rate_hz per second, built
    ahead of time, or given as `built`, the build_request of each spec), then
    collect every result.  Returns per-request records and the wall time from
    the first arrival to the last completion."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, len(specs))
    if built is None:
        built = [build_request(s) for s in specs]
    t0 = time.perf_counter()
    arrival = t0 + np.cumsum(gaps) - gaps[0]
    ids = []
    for k, (spec, (H, es, c, seeded)) in enumerate(zip(specs, built)):
        now = time.perf_counter()
        if arrival[k] > now:
            time.sleep(arrival[k] - now)
        ids.append(eng.submit(H, es, c, seeded=seeded))
        if log:
            log(f"submitted {k}: N={spec.N} b={spec.b} S={spec.S} M={spec.M}")
    out = []
    for k, rid in enumerate(ids):
        r = eng.wait(rid, want_D=want_D)
        out.append((specs[k], r))
    wall = time.perf_counter() - t0
    return out, wall


def summarize(records, wall: float) -> dict:
    lat = np.array([r.total_ms for _, r in records])
    upd = sum(s.N * s.M * s.S for s, _ in records)
    return {"requests": len(records), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "updates_per_s": upd / wall, "wall_s": wall}
