"""fkpm_moments_noalloc: the moments of fkpm_set_H + fkpm_set_R + fkpm_moments
from inputs already in device memory, with no device allocation (working
memory in the library's static scratch).  Checked against the engine's own
fkpm_moments on the same H and R, FP64 at 1e-10, FP32-mixed at the FP32
recursion bound, on two calls in a row (the second one smaller, so that it
runs on what the first left in the scratch); the device's free memory does
not change across the calls, and a problem larger than the scratch is
refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls the test needs for its own buffers."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.lib.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.bufs = []

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
        self.bufs.append(p.value)
        assert self.lib.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p.value

    def zeros(self, nbytes: int) -> int:
        return self.upload(np.zeros(nbytes, dtype=np.uint8))

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes,
                                  D2H) == 0
        return out

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def free_bytes(self) -> int:
        f, t = C.c_size_t(), C.c_size_t()
        assert self.lib.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    def release(self):
        for p in self.bufs:
            self.lib.hipFree(C.c_void_p(p))
        self.bufs = []


def _case(oracle, shape, seed):
    from fastkpm_amd import lattice, EnergyScale
    H = lattice.lattice(shape, b=2, J=1.5, W=2.0, seed=seed)
    groups = lattice.colouring(shape, tuple(4 for _ in shape), 2)
    R = oracle.set_R_correlated(groups, oracle.MT19937(seed + 1))
    return H, np.asfortranarray(R), EnergyScale(-7.0, 7.0)


def _engine_moments(H, R, es, M, prec):
    from fastkpm_amd import HipEngine
    eng = HipEngine("cx_double", prec)
    eng.set_R(R)
    eng.set_H(H, es)
    mu = eng.moments(M)
    eng.close()
    return mu


def _stage(hip, H, R, M):
    """Device copies of H's pattern and values, R and an output for mu."""
    return (hip.upload(H.row_ptr.astype(np.int32)), hip.upload(H.col_idx.astype(np.int32)),
            hip.upload(H.val.astype(np.complex128)),
            hip.upload(R.astype(np.complex128).reshape(-1, order="F")), hip.zeros(8 * M))


def _call(pc, H, R, es, M, bufs):
    from fastkpm_amd import _capi
    rp, ci, hv, rd, mu_d = bufs
    n, s = R.shape
    return _capi.lib().fkpm_moments_noalloc(0, pc, H.n_rows, H.b_len, H.n_blocks(), rp, ci, hv,
                                            es.avg(), es.mag(), n, s, rd, M, mu_d, None)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_moments_noalloc_two_calls(oracle, prec):
    from fastkpm_amd import _capi
    pc = _capi.PREC_FP64 if prec == "fp64" else _capi.PREC_FP32
    tol = FP64_TOL if prec == "fp64" else 2e-5
    hip = Hip()
    try:
        # the second call is smaller and runs on what the first left in the scratch
        cases = [(_case(oracle, (32, 32), 3), 200),   # N = 2048, S = 32, M = 200
                 (_case(oracle, (8, 8), 5), 64)]      # N = 128, S = 32, M = 64
        refs = [_engine_moments(H, R, es, M, prec) for (H, R, es), M in cases]
        staged = [_stage(hip, H, R, M) for (H, R, es), M in cases]
        warm = _stage(hip, cases[1][0][0], cases[1][0][1], 16)
        hip.sync()
        # first use on this device: the scratch's address is looked up
        assert _call(pc, *cases[1][0], 16, warm) == 0, _capi.lib().fkpm_last_error()
        hip.sync()
        free0 = hip.free_bytes()
        rcs = [_call(pc, H, R, es, M, b) for ((H, R, es), M), b in zip(cases, staged)]
        hip.sync()
        assert rcs == [0, 0], _capi.lib().fkpm_last_error()
        # no device memory changed hands inside the two calls
        assert hip.free_bytes() == free0
        for ((H, R, es), M), b, ref in zip(cases, staged, refs):
            mu = hip.download(b[4], np.zeros(M))
            assert rel_err(mu, ref) < tol
            assert mu[0] == pytest.approx(H.n_rows * H.b_len, rel=1e-12)
    finally:
        hip.release()


def test_moments_noalloc_refuses_beyond_scratch(oracle):
    from fastkpm_amd import _capi, lattice, EnergyScale
    hip = Hip()
    try:
        # N = 2^17 orbitals x 64 columns: four FP64 vectors alone take 512 MiB
        H = lattice.lattice((256, 256), b=2, J=1.0, seed=1)
        n = H.n_rows * 2
        R = np.zeros((n, 64), dtype=np.complex128, order="F")
        R[np.arange(n), np.arange(n) % 64] = 1.0
        bufs = _stage(hip, H, R, 16)
        rc = _call(_capi.PREC_FP64, H, R, EnergyScale(-7, 7), 16, bufs)
        assert rc == -1
        assert b"static scratch" in _capi.lib().fkpm_last_error()
    finally:
        hip.release()
