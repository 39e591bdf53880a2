"""A seeded request stream over the ranges of the batched-serving workload
(2-D / 3-D lattices, 2-spin or 8-orbital blocks, N from 4096 to about 1M,
M from 256 to 2048, S from 16 to 128, disorder W in [0, 6]) through the
batched engine, submitted at Poisson arrival times to an HBM arena smaller
than the stream's total footprint, so that admission has to wait for
retirements.  Each request is the energy-and-force solve moments(M) +
autodiff_matrix(c, D) on H's pattern (src/engine_cpu.cpp:63-84, 255-304).

- requests small enough for the oracle (N about 4096-8192, M = 256): moments
  and dE/dH against the oracle at 1e-10 on the same H and the same R (the
  device draws R's phases from the column seeds bit-identically to
  set_R_correlated, src/engine.cpp:76-102);
- the largest requests (N = 1,024,000 and 1,048,352, and one with M = 2048):
  mu_0 = N (probing R has one unit-modulus entry per row), the trace of dE/dH
  against a central finite difference of E under H -> H +- eps I at a fixed
  energy scale (both shifted solves run through the same stream), and the
  FP32-mixed energy within 1e-6 of FP64;
- every request's moments and D are bitwise the same when it runs alone on
  the idle engine as in the stream.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
EPS = 1e-3


def _specs():
    from fastkpm_amd.serving import RequestSpec
    rng = np.random.default_rng(5)
    W = lambda: float(rng.uniform(0.0, 6.0))
    seed = lambda: int(rng.integers(1 << 30))
    return {
        # oracle-sized: the smallest N of each kind
        "small_2d_b2": RequestSpec(2, 2, 46, 4, 256, W(), seed()),     # N 4232, S 32
        "small_2d_b8": RequestSpec(2, 8, 24, 2, 256, W(), seed()),     # N 4608, S 32
        "small_3d_b2": RequestSpec(3, 2, 16, 2, 256, W(), seed()),     # N 8192, S 16
        # the largest N of the ranges, and the largest M
        "big_3d_b2": RequestSpec(3, 2, 80, 4, 256, W(), seed()),       # N 1024000, S 128
        "big_2d_b8": RequestSpec(2, 8, 362, 2, 256, W(), seed()),      # N 1048352, S 32
        "longM_2d_b2": RequestSpec(2, 2, 128, 8, 2048, W(), seed()),   # N 32768, S 128, M 2048
        "mid_3d_b8": RequestSpec(3, 8, 16, 2, 512, W(), seed()),       # N 32768, S 64
    }


SMALL = ("small_2d_b2", "small_2d_b8", "small_3d_b2")
BIG = ("big_3d_b2", "big_2d_b8", "longM_2d_b2")


def _shift(H, eps):
    """H + eps I (the diagonal of every diagonal block)."""
    H2 = H.like()
    H2.val[:] = H.val
    b = H.b_len
    rows = np.repeat(np.arange(H.n_rows), np.diff(H.row_ptr))
    diag_p = np.flatnonzero(rows == np.asarray(H.col_idx))
    blk = H2.val.reshape(-1, b, b)
    blk[diag_p] += eps * np.eye(b)[None]
    return H2


def _trace(D, H) -> float:
    b = H.b_len
    rows = np.repeat(np.arange(H.n_rows), np.diff(H.row_ptr))
    diag_p = np.flatnonzero(rows == np.asarray(H.col_idx))
    blk = D.reshape(-1, b, b)[diag_p]
    return float(np.sum(np.real(np.diagonal(blk, axis1=1, axis2=2))))


@pytest.fixture(scope="module")
def stream():
    from fastkpm_amd import serving
    specs = _specs()
    built = {k: serving.build_request(s) for k, s in specs.items()}
    jobs = []  # (name, H, es, c, seeded, want_D)
    for k in specs:
        H, es, c, sd = built[k]
        jobs.append((k, H, es, c, sd, True))
        if k in BIG:
            for sgn in (+1, -1):
                jobs.append((f"{k}{'+' if sgn > 0 else '-'}", _shift(H, sgn * EPS), es, c, sd,
                             False))
    fp = {j[0]: serving.footprint("fp64", j[1], len(j[4][1]), len(j[3])) for j in jobs}
    total = sum(fp.values())
    # the arena holds the largest request and a little more, far less than all
    arena = int(max(fp.values()) * 1.3)
    assert arena < total / 2
    eng = serving.BatchEngine("fp64", hbm_bytes=arena, max_slots=16)
    rng = np.random.default_rng(11)
    gaps = rng.exponential(1.0 / 20.0, len(jobs))  # Poisson arrivals, 20 per second
    import time
    ids = {}
    t0 = time.perf_counter()
    t_next = t0
    for (name, H, es, c, sd, wd), gap in zip(jobs, gaps):
        t_next += gap
        dt = t_next - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids[name] = (eng.submit(H, es, c, seeded=sd), wd)
    res = {name: eng.wait(rid, want_D=wd) for name, (rid, wd) in ids.items()}
    st = eng.stats()
    yield dict(specs=specs, built=built, res=res, eng=eng, stats=st, fp=fp, arena=arena)
    eng.close()


def test_stream_admission_waited_for_memory(stream):
    """The arena could not hold the stream at once: the engine ran it in
    several slot sets, never above the arena."""
    st = stream["stats"]
    assert st["peak_bytes"] <= stream["arena"]
    assert st["captures"] >= 2 and st["peak_slots"] < len(stream["res"])


@pytest.mark.parametrize("name", SMALL)
def test_stream_small_requests_match_oracle(stream, oracle, name):
    from fastkpm_amd import host
    spec = stream["specs"][name]
    H, es, c, (groups, seeds) = stream["built"][name]
    r = stream["res"][name]
    R = host.set_R_correlated("cx_double", groups, host.RNG(spec.seed))
    orc = oracle.Engine()
    orc.set_H(H, es.lo, es.hi)
    orc.set_R(R)
    assert rel_err(r.mu, orc.moments(spec.M)) < FP64_TOL
    assert r.energy == pytest.approx(float(np.dot(c, r.mu)), rel=1e-12)
    assert rel_err(r.D, orc.autodiff_matrix(c, H)) < FP64_TOL


@pytest.mark.parametrize("name", BIG)
def test_stream_large_requests_properties(stream, name):
    spec = stream["specs"][name]
    H = stream["built"][name][0]
    r = stream["res"][name]
    N = H.n_rows * H.b_len
    assert N == spec.N and len(r.mu) == spec.M
    assert r.mu[0] == pytest.approx(N, rel=1e-13)
    e_p = stream["res"][name + "+"].energy
    e_m = stream["res"][name + "-"].energy
    fd = (e_p - e_m) / (2 * EPS)
    tr = _trace(r.D, H)
    assert abs(tr - fd) < 1e-6 * abs(fd), (tr, fd)


def test_stream_fp32_energies_within_1e6(stream):
    from fastkpm_amd import serving
    eng = serving.BatchEngine("fp32", hbm_bytes=stream["arena"], max_slots=16)
    try:
        ids = {}
        for name in BIG + ("mid_3d_b8",):
            H, es, c, sd = stream["built"][name]
            ids[name] = eng.submit(H, es, c, seeded=sd)
        for name, rid in ids.items():
            e32 = eng.wait(rid, want_D=False).energy
            e64 = stream["res"][name].energy
            assert abs(e32 - e64) < 1e-6 * abs(e64), (name, e32, e64)
    finally:
        eng.close()


def test_stream_results_same_alone(stream):
    """Each request submitted alone to the idle engine: bitwise the same."""
    eng = stream["eng"]
    for name in stream["specs"]:
        H, es, c, sd = stream["built"][name]
        alone = eng.wait(eng.submit(H, es, c, seeded=sd))
        mixed = stream["res"][name]
        assert np.array_equal(alone.mu, mixed.mu), name
        assert np.array_equal(alone.D, mixed.D), name
