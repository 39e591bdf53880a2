#!/usr/bin/env python3
"""Benchmark of the MI355X stochastic Chebyshev engine.

One step = one full KPM solve of the workload with H, R and the coefficients
resident in HBM:
  moments(M) -> E = sum c_m mu_m,
  autodiff_matrix(c_E, D) (dE/dH on H's pattern),
  stoch_matrix(c_f, D')   (D_ij = f(H)_ij on H's pattern),
each D copied to the host as the drop-in Engine calls return it.
value = N*M*S / (seconds per step): Chebyshev vector updates per second.

Default workload (the largest one-GPU configuration, and the one the
FP32-mixed vs FP64 comparison is stated on): cubic lattice 64^3 sites x 2 spin
orbitals (N = 524288, complex 2x2 BSR blocks), hopping -I2, Anderson on-site
disorder U[-2, 2] and Kondo coupling -J s_i.sigma with J = 1 and seeded
random unit spins, M = 2000, S = 128 probing colours (4x4x4 spatial colouring
x 2 orbitals), Fermi function at mu = 0, kT = 0.01; FP64, with the FP32-mixed
run of the same workload beside it.

Prints one JSON line.  Besides the headline it carries: "fp32_mixed" (the same
steps in FP32 storage with FP64 accumulation, errors against FP64); "serving"
(a seeded Poisson stream of mixed energy + dE/dH requests over the full
serving ranges through the batched engine, FP64 and FP32-mixed: p50/p99
latency and aggregate updates/s; the p99 is also in "latency_ms");
"C2" (the 2-D Kondo lattice, FP64 and FP32-mixed); "C1" (the 32x32 lattice
solve time, which per-step launches set); "conductivity" (two-current moments
moments2_v1(M=128) of the x current on C2 and the static sigma_xx);
"exact_check" (KPM against exact diagonalisation on a small lattice of the
same kind); "cpu_baseline" (the FP64 oracle on the same H and R with M'
moments, a bounded sample, one thread and on column shards).
--serve-only runs the serving stream alone and prints it as its own line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from fastkpm_amd import HipEngine, bandwidth_probe, device_count, host, lattice  # noqa: E402
from fastkpm_amd import serving  # noqa: E402
from fastkpm_amd.engine import EnergyScale  # noqa: E402

METRIC = ("N·M·S Chebyshev updates/s, FP32-mixed vs FP64 at ≤1e-6 energy err; "
          "p99 lat")
HBM_PEAK_GBS = 8000.0


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


WORKLOADS = {
    "C1": "square lattice 32x32 periodic, t=-1, real H, N=1024, es=[-4.1,4.1], M=500, "
          "S=16 probing colours (4x4), fermi mu=0 kT=0",
    "C2": "Kondo lattice 256x256 sites x 2 spin orbitals, complex 2x2 BSR, t=-1, J=2, "
          "M=1000, S=64 (4x8 colours x 2 spins), kT=0.01",
    "C3": "cubic 64^3 sites x 2 spin orbitals, Anderson W=4 + Kondo J=1, complex 2x2 BSR, "
          "M=2000, S=128 (4x4x4 colours x 2), kT=0.01",
    "C4": "cubic 32^3 sites with 8x8 orbital blocks (random Hermitian on-site, complex "
          "Gaussian(0,0.5) hopping), M=1024, S=256 (4x4x2 colours x 8), kT=0.01",
}


def build_workload(name: str, seed: int):
    H, groups, cfg = lattice.make_config(name, seed=seed)
    return H, groups, cfg


def algorithmic_bytes(H, N: int, S: int, v: int, vacc: int = 16, table_K: int | None = None):
    """Algorithmic HBM bytes per launch of each step kernel (S logical columns,
    v bytes per vector element, vacc bytes per FP64 accumulator element).
    table_K: stoch_matrix ran from the kept forward vectors (one pass over
    alpha_0 .. alpha_K, then K Chebyshev steps), so its figure is the call's
    bytes over its K + 1 launches."""
    b = H.b_len
    nnzb = H.n_blocks()
    h_bytes = nnzb * b * b * v + 4 * nnzb + 4 * (H.n_rows + 1)
    vec = N * S * v
    acc = N * S * vacc
    d_bytes = nnzb * b * b * vacc
    return {
        # read alpha_m, alpha_{m-1}, write alpha_{m+1}, read H
        "moments": 3 * vec + h_bytes,
        # Clenshaw step b_k = c_k R + 2 H b_{k+1} - b_{k+2}: read b_{k+1}, b_{k+2},
        # write b_k, read H (R in compact form: 4 + v bytes per orbital row)
        "stoch": (3 * vec + h_bytes + N * (4 + v) if table_K is None else
                  # read alpha_0..alpha_K, write u and v; K steps; read v once more
                  ((table_K + 1) * vec + 2 * vec + table_K * (3 * vec + h_bytes) + vec)
                  / (table_K + 1)),
        # read alpha_m, alpha_{m+1}, beta_{m+1}, beta_{m+2}; write alpha_{m-1},
        # beta_m; read H; read-modify-write D (FP64)
        "autodiff": 6 * vec + h_bytes + 2 * d_bytes,
    }


def measured_traffic(config: str, prec: str, kernel: str, kc: dict, b: int):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE measurement of this workload
    (profiles/traffic_<config>_<prec>.json): the entry of the same kernel,
    storage type, block size and geometry at the same grid, or None."""
    path = ROOT / "profiles" / f"traffic_{config.lower()}_{prec}.json"
    if not path.exists():
        return None
    data = json.loads(path.read_text())
    st = "double" if prec == "fp64" else "float"
    stem = f"::{kernel.split(' ')[0]}<{st}, true, {b}, {kc['lanes']}, {kc['kq']}"
    for k in data.get("kernels", []):
        if stem in k["kernel"] and k["grid_threads"] == 256 * kc["n_wg"]:
            out = {"traffic_bytes_per_launch": k["traffic_bytes_per_launch"],
                   "source": f"{path.name} (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)"}
            for key in ("request_bytes_per_launch", "read_request_bytes_per_launch"):
                if k.get(key):
                    out[key] = k[key]
            return out
    return None


def run_gpu(H, groups, cfg, prec: str, steps: int, warmup: int, seed: int, tune: bool = True):
    M = cfg["M"]
    eng = HipEngine("cx_double", prec)
    t_es = time.perf_counter()
    es = EnergyScale(*cfg["es"]) if "es" in cfg else eng.energy_scale(H, 0.01)
    t_es = time.perf_counter() - t_es
    eng.set_R_correlated(groups, host.RNG(seed))
    eng.set_H(H, es)
    t_tune = time.perf_counter()
    if tune:
        eng.autotune()  # setup, like a library's heuristics pass: not in the timed region
    t_tune = time.perf_counter() - t_tune
    S = eng.R.shape[1]
    N = H.n_rows * H.b_len
    Mq = 2 * M
    cE = host.expansion_coefficients(M, Mq, host.FERMI_ENERGY, cfg["kT"], cfg["mu"], es.lo, es.hi)
    cF = host.expansion_coefficients(M, Mq, host.FERMI_DENSITY, cfg["kT"], cfg["mu"], es.lo, es.hi)
    Dg = H.like()
    Ds = H.like()

    def step():
        # as Engine::autodiff_matrix / stoch_matrix: each D ends in host memory
        mu = eng.moments(M)
        eng.autodiff_matrix(cE, Dg)
        eng.stoch_matrix(cF, Ds)
        return mu

    for _ in range(warmup):
        step()
    eng.set_timing(True)
    tot = {"moments": [0.0, 0], "autodiff": [0.0, 0], "stoch": [0.0, 0]}
    times = []
    eng.synchronize()
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        mu = eng.moments(M)
        tm = eng.timing()
        eng.autodiff_matrix(cE, Dg)
        ta = eng.timing()
        eng.stoch_matrix(cF, Ds)
        ts = eng.timing()
        times.append(time.perf_counter() - t0)
        tot["moments"][0] += tm["moments_ms"]
        tot["moments"][1] += tm["moments_launches"]
        tot["autodiff"][0] += ta["autodiff_ms"]
        tot["autodiff"][1] += ta["autodiff_launches"]
        tot["stoch"][0] += ts["stoch_ms"]
        tot["stoch"][1] += ts["stoch_launches"]
    eng.synchronize()
    wall = time.perf_counter() - t_all
    eng.set_timing(False)
    stoch_method = eng.stoch_method()  # from the kept forward vectors, or Clenshaw
    energy = float(np.dot(cE, mu))
    # the device-to-host copy of one D (pageable host memory, as a caller's
    # SpMatBsr holds it): two of them are inside every timed step
    t_d2h = time.perf_counter()
    for _ in range(3):
        eng.get_D(Ds)
    d2h_ms = 1e3 * (time.perf_counter() - t_d2h) / 3
    eng.moments(M)
    alpha_table = eng.alpha_table()  # autodiff reads the kept forward vectors
    eng.autodiff_matrix(cE, Dg)
    eng.stoch_matrix(cF, Ds)
    kcfg = {name: eng.kernel_config(k) for k, name in
            ((0, "moments"), (1, "xi"), (2, "autodiff"), (3, "outer"), (4, "moments_table"),
             (5, "autodiff_table"))}
    return dict(eng=eng, es=es, N=N, S=S, M=M, mu=mu, energy=energy, wall=wall, times=times,
                kernel_ms=tot, Dg=Dg.val.copy(), Ds=Ds.val.copy(), config=eng.config(),
                kernel_config=kcfg, tune_s=t_tune, state=eng.state(), alpha_table=alpha_table,
                stoch_method=stoch_method, d2h_ms=d2h_ms,
                energy_scale_s=None if "es" in cfg else t_es)


SMALL_SHAPES = {"C1": (32, 32), "C2": (16, 16), "C3": (8, 8, 8), "C4": (4, 4, 4)}


def _fermi_energy(x, kT, mu):
    """fermi_energy of src/fastkpm.cpp:375-383, elementwise."""
    x = np.asarray(x, dtype=float)
    if kT < 1e-15:
        return np.where(x < mu, x - mu, 0.0)
    a = (x - mu) / abs(kT)
    soft = -kT * np.log1p(np.exp(-np.clip(a, -20, 20)))
    return np.where(np.abs(a) > 20, np.where(x < mu, x - mu, 0.0), soft)


def _fermi_density(x, kT, mu):
    """fermi_density of src/fastkpm.cpp:385-393, elementwise."""
    x = np.asarray(x, dtype=float)
    if kT < 1e-15:
        return np.where(x < mu, 1.0, 0.0)
    a = (x - mu) / abs(kT)
    soft = 1.0 / (np.exp(np.clip(a, -20, 20)) + 1.0)
    return np.where(np.abs(a) > 20, np.where(x < mu, 1.0, 0.0), soft)


def exact_check(name: str, prec: str, seed: int) -> dict:
    """KPM on the GPU against exact FP64 diagonalisation of f(H) on a small
    lattice of the workload's family, R = identity (no stochastic error): the
    remaining difference is the Chebyshev truncation (Jackson kernel, M)."""
    cfg = lattice.CONFIGS[name]
    H, _, _ = lattice.make_config(name, seed=seed, scale_shape=SMALL_SHAPES[name])
    n = H.n_rows * H.b_len
    M = cfg["M"]
    eng = HipEngine("cx_double", prec)
    t_es = time.perf_counter()
    es = EnergyScale(*cfg["es"]) if "es" in cfg else eng.energy_scale(H, 0.01)
    t_es = time.perf_counter() - t_es
    eng.set_R(np.eye(n, dtype=np.complex128))
    eng.set_H(H, es)
    cE = host.expansion_coefficients(M, 2 * M, host.FERMI_ENERGY, cfg["kT"], cfg["mu"], es.lo,
                                     es.hi)
    cF = host.expansion_coefficients(M, 2 * M, host.FERMI_DENSITY, cfg["kT"], cfg["mu"], es.lo,
                                     es.hi)
    e_kpm = float(np.dot(cE, eng.moments(M)))
    Dg, Ds = H.like(), H.like()
    eng.autodiff_matrix(cE, Dg)
    eng.stoch_matrix(cF, Ds)
    eng.close()
    lam, V = np.linalg.eigh(H.to_dense())
    e_ex = float(np.sum(_fermi_energy(lam, cfg["kT"], cfg["mu"])))
    F = (V * _fermi_density(lam, cfg["kT"], cfg["mu"])) @ V.conj().T  # = dE/dH = f(H)
    mask = np.abs(H.to_dense()) > 0
    fm = F[mask]
    return {"lattice": "x".join(map(str, SMALL_SHAPES[name])) + f" x {H.b_len} orbitals",
            "N": n, "M": M, "precision": prec,
            "energy_rel_err": abs(e_kpm - e_ex) / abs(e_ex),
            "dEdH_rel_err": float(np.max(np.abs(Dg.to_dense()[mask] - fm)) / np.max(np.abs(fm))),
            "Dij_rel_err": float(np.max(np.abs(Ds.to_dense()[mask] - fm)) / np.max(np.abs(fm)))}


def cpu_baseline_threads(H, groups, cfg, es, seed: int, m_sample: int, threads: int):
    """The oracle on S-column shards, one thread per shard (the split of the
    reference's Engine_Threaded): moments, D summed over shards."""
    import concurrent.futures as cf
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_py
    R = host.set_R_correlated("cx_double", groups, host.RNG(seed))
    S = R.shape[1]
    N = H.n_rows * H.b_len
    cE = host.expansion_coefficients(m_sample, 2 * m_sample, host.FERMI_ENERGY, cfg["kT"],
                                     cfg["mu"], es.lo, es.hi)
    cF = host.expansion_coefficients(m_sample, 2 * m_sample, host.FERMI_DENSITY, cfg["kT"],
                                     cfg["mu"], es.lo, es.hi)
    cuts = np.linspace(0, S, threads + 1).astype(int)
    shards = [R[:, a:b] for a, b in zip(cuts[:-1], cuts[1:]) if b > a]

    def work(Rs):
        orc = oracle_py.Engine()
        orc.set_H(H, es.lo, es.hi)
        orc.set_R(np.asfortranarray(Rs))
        D = H.like()
        orc.moments(m_sample)
        orc.autodiff_matrix(cE, D)
        orc.stoch_matrix(cF, D)

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(len(shards)) as ex:
        list(ex.map(work, shards))
    dt = time.perf_counter() - t0
    return {"value": N * m_sample * S / dt, "unit": "updates/s", "cores": len(shards),
            "kind": "port",
            "sample": (f"same workload and M'={m_sample} as the one-thread figure, R split into "
                       f"{len(shards)} column shards run concurrently (oracle calls release the "
                       f"GIL) in {dt:.2f} s")}


def cpu_baseline(H, groups, cfg, es, seed: int, m_sample: int):
    """The FP64 oracle (the reference CPU algorithm restated in C, one thread)
    on the same H, R and coefficients with M' = m_sample moments."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_py
    R = host.set_R_correlated("cx_double", groups, host.RNG(seed))
    S = R.shape[1]
    N = H.n_rows * H.b_len
    cE = host.expansion_coefficients(m_sample, 2 * m_sample, host.FERMI_ENERGY, cfg["kT"],
                                     cfg["mu"], es.lo, es.hi)
    cF = host.expansion_coefficients(m_sample, 2 * m_sample, host.FERMI_DENSITY, cfg["kT"],
                                     cfg["mu"], es.lo, es.hi)
    orc = oracle_py.Engine()
    orc.set_H(H, es.lo, es.hi)
    orc.set_R(R)
    D = H.like()
    t0 = time.perf_counter()
    orc.moments(m_sample)
    orc.autodiff_matrix(cE, D)
    orc.stoch_matrix(cF, D)
    dt = time.perf_counter() - t0
    return {"value": N * m_sample * S / dt, "unit": "updates/s", "cores": 1, "kind": "port",
            "sample": (f"same workload (N={N}, S={S}) with M'={m_sample} moments: one "
                       f"moments+autodiff_matrix+stoch_matrix pass of the FP64 oracle "
                       f"(oracle/oracle.c, one thread) in {dt:.2f} s; host has "
                       f"{os.cpu_count()} cores")}


def single_engine_rate(specs, built, prec: str) -> dict:
    """N*M*S per second of the requests solved one by one on HipEngine
    (moments + autodiff_matrix, D to the host) with H and R set before each
    timed solve and the untuned default kernels, as the batched engine runs
    them."""
    t_tot, upd = 0.0, 0
    for sp, (H, es, c, (groups, seeds)) in zip(specs, built):
        eng = HipEngine("cx_double", prec)
        eng.set_R_correlated(groups, host.RNG(sp.seed))
        eng.set_H(H, es)
        D = H.like()
        eng.moments(sp.M)  # first use of the shape's buffers
        eng.autodiff_matrix(c, D)
        eng.synchronize()
        t0 = time.perf_counter()
        eng.moments(sp.M)
        eng.autodiff_matrix(c, D)
        t_tot += time.perf_counter() - t0
        upd += sp.N * sp.M * sp.S
        eng.close()
    return {"updates_per_s": upd / t_tot, "solve_s_total": t_tot}


def serving_leg(n_req: int, n_max: int, m_max: int, rate: float, seed: int, hbm_gb: int,
                precs=("fp64", "fp32"), graph: bool = True) -> dict:
    """The mixed Poisson request stream through the batched engine."""
    rng = np.random.default_rng(seed)
    specs = [serving.draw_spec(rng, n_max=n_max, m_max=m_max) for _ in range(n_req)]
    out = {"stream": (f"{n_req} requests, Poisson arrivals at {rate}/s (seed {seed}), then the "
                      f"same requests queued at once ('saturated'); 2-D/3-D "
                      f"lattices, 2-spin or 8-orbital blocks, N 4096..{n_max}, M 256..{m_max}, "
                      f"S 16..128, W in [0,6]; each request = moments + autodiff_matrix "
                      f"(energy and dE/dH), HBM arena {hbm_gb} GiB"),
           "updates_total": int(sum(sp.N * sp.M * sp.S for sp in specs))}
    built = [serving.build_request(sp) for sp in specs]  # host-side H and coefficients, once
    for prec in precs:
        eng = serving.BatchEngine(prec, hbm_bytes=hbm_gb << 30, max_slots=64, graph=graph)
        # first use: kernel loading and the first graph capture
        warm = serving.draw_spec(np.random.default_rng(seed + 1), n_max=8192, m_max=256)
        H, es, c, sd = serving.build_request(warm)
        eng.wait(eng.submit(H, es, c, seeded=sd), want_D=False)
        recs, wall = serving.run_stream(eng, specs, rate, seed=seed, built=built)
        summ = serving.summarize(recs, wall)
        summ["energies"] = [r.energy for _, r in recs]
        # capacity: the same requests all queued at once
        recs2, wall2 = serving.run_stream(eng, specs, 1e9, seed=seed, built=built)
        sat = serving.summarize(recs2, wall2)
        summ["saturated"] = {"updates_per_s": sat["updates_per_s"], "wall_s": sat["wall_s"],
                             "p50_ms": sat["p50_ms"], "p99_ms": sat["p99_ms"]}
        summ.update(eng.stats())
        out[prec] = summ
        eng.close()
        # the same work on the single-request engine, one request after
        # another (moments + autodiff_matrix with H and R already set): the
        # rate the batched engine is held to
        summ["single_engine"] = single_engine_rate(specs, built, prec)
        summ["saturated_vs_single_engine"] = (summ["saturated"]["updates_per_s"]
                                              / summ["single_engine"]["updates_per_s"])
    if "fp64" in out and "fp32" in out:
        e64 = np.array(out["fp64"].pop("energies"))
        e32 = np.array(out["fp32"].pop("energies"))
        out["fp32_energy_rel_err_max"] = float(np.max(np.abs(e32 - e64) / np.abs(e64)))
    for prec in precs:
        out[prec].pop("energies", None)
    return out


def conductivity_leg(H, groups, cfg, es, seed: int, M2: int, reps: int = 2) -> dict:
    """Two-current moments (moments2_v1) of the x current on the same workload
    and the static conductivity sigma_xx from them, FP64 and FP32-mixed."""
    shape = cfg["shape"]
    jx = lattice.current(H, shape, 0)
    N = H.n_rows * H.b_len
    c2 = host.conductivity_coefficients(M2, 2 * M2, max(cfg.get("kT", 0.01), 1e-3),
                                        cfg.get("mu", 0.0), 0.0, es.lo, es.hi)
    out = {"M": M2, "current": "j_x = i[H, X] (bonds along x)", "reps": reps}
    res = {}
    for prec in ("fp64", "fp32"):
        eng = HipEngine("cx_double", prec)
        eng.set_R_correlated(groups, host.RNG(seed))
        eng.set_H(H, es)
        S = eng.R.shape[1]
        eng.moments2_v1(M2, jx, jx)  # warm-up: buffers
        eng.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            mu2 = eng.moments2_v1(M2, jx, jx)
        t = (time.perf_counter() - t0) / reps
        eng.close()
        # complex GEMM: 8 real flops per multiply-add, M^2 N S of them; the
        # recursions: 2 M products with H and M with j2, one with j1
        gemm_flops = 8.0 * M2 * M2 * N * S
        res[prec] = mu2
        out[prec] = {"ms": round(1e3 * t, 2), "sigma_xx": complex(np.sum(c2 * mu2)).real,
                     "gemm_TFLOPs_if_all_time": round(gemm_flops / t / 1e12, 2),
                     "updates_per_s": 2 * M2 * N * S / t}
    out["fp32_mu2_rel_err_vs_fp64"] = float(np.max(np.abs(res["fp32"] - res["fp64"]))
                                            / np.max(np.abs(res["fp64"])))
    return out


def small_solve_leg(name: str, seed: int, reps: int = 20) -> dict:
    """Wall time of one full solve (moments + autodiff_matrix + stoch_matrix,
    both D's to the host) of a workload too small to fill the GPU (C1: 1024
    orbitals x 16 probe columns), where the per-step launches, not HBM, set
    the time: compared with the sum of the step kernels' device time."""
    H, groups, cfg = build_workload(name, seed)
    r = run_gpu(H, groups, cfg, "fp64", reps, 2, seed, tune=False)
    wall_ms = 1e3 * r["wall"] / reps
    launches = sum(n for _, n in r["kernel_ms"].values()) // reps
    dev_ms = sum(ms for ms, _ in r["kernel_ms"].values()) / reps
    return {"workload": f"{name}: {WORKLOADS[name]}", "N": r["N"], "S": r["S"], "M": r["M"],
            "solve_ms": round(wall_ms, 3), "step_launches_per_solve": int(launches),
            "step_loops_device_ms": round(dev_ms, 3),
            "us_per_step_launch": round(1e3 * dev_ms / max(launches, 1), 2),
            "energy": r["energy"]}


def extra_config_leg(name: str, steps: int, warmup: int, seed: int) -> dict:
    """FP64 and FP32-mixed steps of another workload (fewer steps than the
    headline), with per-kernel roofline figures."""
    H, groups, cfg = build_workload(name, seed)
    out = {"workload": f"{name}: {WORKLOADS[name]}", "N": H.n_rows * H.b_len, "M": cfg["M"],
           "steps": steps}
    res = {}
    for prec in ("fp64", "fp32"):
        r = run_gpu(H, groups, cfg, prec, steps, warmup, seed)
        N, S, M = r["N"], r["S"], r["M"]
        v = 16 if prec == "fp64" else 8
        ab = algorithmic_bytes(H, N, S, v,
                               table_K=(M + 1) // 2 if r["stoch_method"] == "table" else None)
        kern = {}
        for k, (ms, n) in r["kernel_ms"].items():
            if n:
                us = 1e3 * ms / n
                kern[k] = {"avg_launch_us": round(us, 3),
                           "frac": round(ab[k] / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
        res[prec] = r
        out[prec] = {"ms_per_step": round(1e3 * r["wall"] / steps, 3),
                     "value": N * M * S / (r["wall"] / steps), "kernels": kern,
                     "stoch_method": r["stoch_method"], "alpha_table": r["alpha_table"],
                     "d2h_ms_per_D": round(r["d2h_ms"], 3)}
        out["S"] = S
    out["fp32_speedup_vs_fp64"] = round(res["fp64"]["wall"] / res["fp32"]["wall"], 3)
    out["fp32_energy_rel_err_vs_fp64"] = abs(res["fp32"]["energy"] - res["fp64"]["energy"]) / abs(
        res["fp64"]["energy"])
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C3")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-sample-moments", type=int, default=0,
                    help="M' of the CPU baseline (0: about 5e8 updates of the workload)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the C2 and C1 legs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-mixed comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-conductivity", action="store_true",
                    help="skip the two-current (conductivity) moments")
    ap.add_argument("--conductivity-moments", type=int, default=128)
    ap.add_argument("--conductivity-only", action="store_true",
                    help="run only the two-current moments leg and print it as one JSON line")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving stream, at the full request ranges")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--serve-hbm-gb", type=int, default=64)
    ap.add_argument("--serve-no-graph", action="store_true",
                    help="serving: launch every tick directly instead of replaying hipGraphs")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the engine runs on a single device "
              "(S-column sharding over several GPUs is designed, not built); rerun with --gpus 1",
              file=sys.stderr)
        return 2

    if device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1

    if args.serve_only:
        _log("serving stream (full ranges)")
        sv = serving_leg(args.serve_requests, 1 << 20, 2048, args.serve_rate, args.seed,
                         args.serve_hbm_gb, graph=not args.serve_no_graph)
        print(json.dumps({"metric": "batched serving: aggregate N·M·S updates/s and p50/p99 "
                                    "latency of a mixed KPM request stream",
                          "value": sv["fp64"]["updates_per_s"], "unit": "updates/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic", "serving": sv}), flush=True)
        return 0

    if args.conductivity_only:
        H, groups, cfg = build_workload(args.config, args.seed)
        eng = HipEngine("cx_double", "fp64")
        es = EnergyScale(*cfg["es"]) if "es" in cfg else eng.energy_scale(H, 0.01)
        eng.close()
        _log(f"two-current moments, M={args.conductivity_moments}")
        cv = conductivity_leg(H, groups, cfg, es, args.seed, args.conductivity_moments)
        N = H.n_rows * H.b_len
        print(json.dumps({"metric": "two-current (conductivity) moments moments2_v1: ms per call",
                          "value": cv["fp64"]["ms"], "unit": "ms", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": f"{args.config}: "
                                                 f"{WORKLOADS.get(args.config, args.config)}",
                                     "N": N, "M": args.conductivity_moments},
                          "conductivity": cv}), flush=True)
        return 0

    # the practical streaming ceiling on this device: a plain 1 GiB copy
    # (k_copy16); its dispatches also calibrate the HBM byte counters of a
    # profiled run of this script
    probe = HipEngine("cx_double", "fp64")
    copy_gbs = bandwidth_probe(probe, 1 << 30, 5)
    probe.close()
    _log(f"building workload {args.config}")
    H, groups, cfg = build_workload(args.config, args.seed)
    _log(f"N={H.n_rows * H.b_len} nnzb={H.n_blocks()} M={cfg['M']}; {args.precision} run")
    r = run_gpu(H, groups, cfg, args.precision, args.steps, args.warmup, args.seed)
    N, S, M = r["N"], r["S"], r["M"]
    ms_per_step = 1e3 * r["wall"] / args.steps
    value = N * M * S / (r["wall"] / args.steps)
    v = 16 if args.precision == "fp64" else 8
    abytes = algorithmic_bytes(H, N, S, v,
                               table_K=(M + 1) // 2 if r["stoch_method"] == "table" else None)
    kern = {}
    for k, (ms, n) in r["kernel_ms"].items():
        if n == 0:
            continue
        avg_us = 1e3 * ms / n
        gbs = abytes[k] / (avg_us * 1e-6) / 1e9
        kern[k] = {"launches_per_step": n // args.steps, "avg_launch_us": round(avg_us, 3),
                   "bytes_per_launch": int(abytes[k]), "achieved_GBs": round(gbs, 1),
                   "frac": round(gbs / HBM_PEAK_GBS, 4),
                   "ms_per_step": round(ms / args.steps, 3)}
    dom = max(kern, key=lambda k: kern[k]["ms_per_step"])
    # the forward and backward steps run the table kinds when the forward
    # vectors are kept; the Clenshaw steps always the ring kind
    kind = {"moments": "moments", "autodiff": "autodiff", "stoch": "moments"}[dom]
    if r["alpha_table"] and dom != "stoch":
        kind += "_table"
    kc = r["kernel_config"][kind]
    if kc.get("mfma"):
        rr = "_rr" if kc.get("mfma_variant") == 2 else ""
        names = {"moments": f"k_mfma_step{rr}", "stoch": f"k_mfma_step{rr} (Clenshaw)",
                 "autodiff": f"k_mfma_autodiff{rr}"}
        geo = f"b={H.b_len}, matrix cores, wg={kc['n_wg']}"
    else:
        names = {"moments": "k_step_moments",
                 "stoch": ("k_table_sums + k_step_moments (from the kept forward vectors)"
                           if r["stoch_method"] == "table" else "k_step_moments (Clenshaw)"),
                 "autodiff": "k_step_autodiff_stored" if r["alpha_table"] else "k_step_autodiff"}
        geo = f"b={H.b_len}, lanes={kc['lanes']}, kq={kc['kq']}, wg={kc['n_wg']}, nt={kc['nt']}"
    roof = {"bound": "hbm", "achieved": kern[dom]["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": kern[dom]["frac"], "traffic": None,
            "kernel": f"{names[dom]} ({geo})",
            "bytes_per_launch": kern[dom]["bytes_per_launch"],
            "avg_launch_us": kern[dom]["avg_launch_us"]}
    # arithmetic of the block products: 8 b^2 real flops per block and column
    # (one SpMM in moments / stoch, two SpMMs and the outer product in autodiff)
    fl = {"moments": 1, "stoch": 1, "autodiff": 3}
    for k in kern:
        flops = fl[k] * H.n_blocks() * S * 8 * H.b_len ** 2
        kern[k]["flops_per_launch"] = flops
        kern[k]["achieved_TFLOPs"] = round(flops / (kern[k]["avg_launch_us"] * 1e-6) / 1e12, 2)
    if kc.get("mfma"):
        peak_tf = 78.6 if args.precision == "fp64" else 157.3
        roof["mfma"] = {"achieved": kern[dom]["achieved_TFLOPs"], "peak": peak_tf,
                        "unit": "TFLOP/s",
                        "frac": round(kern[dom]["achieved_TFLOPs"] / peak_tf, 4)}
    roof["copy_GBs"] = round(copy_gbs, 1)
    roof["frac_of_copy"] = round(kern[dom]["achieved_GBs"] / copy_gbs, 4)
    tr = None if kc.get("mfma") else measured_traffic(args.config, args.precision, names[dom],
                                                        kc, H.b_len)
    if tr is not None:
        roof["traffic"] = tr["traffic_bytes_per_launch"]
        roof["traffic_source"] = tr["source"]
        if "request_bytes_per_launch" in tr:
            # the same bytes from the EA request counters by size (32/64/128 B)
            roof["traffic_by_request_size"] = tr["request_bytes_per_launch"]
        if "read_request_bytes_per_launch" in tr:
            roof["read_bytes_by_request_size"] = tr["read_request_bytes_per_launch"]
    times_ms = np.array(r["times"]) * 1e3
    out = {
        "metric": METRIC, "value": value, "unit": "updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": (f"{args.config}: {WORKLOADS.get(args.config, args.config)}; "
                                f"step = moments + autodiff_matrix + stoch_matrix"),
                   "N": N, "M": M, "S": S, "b": H.b_len, "nnzb": H.n_blocks(),
                   "precision": args.precision, "engine_T": "cx_double"},
        "latency_ms": {"p50": round(float(np.percentile(times_ms, 50)), 3),
                       "p99": round(float(np.percentile(times_ms, 99)), 3)},
        "roofline": roof,
        "kernels": kern,
        "kernel_config": r["kernel_config"],
        "stoch_method": r["stoch_method"],
        "autotune_s": round(r["tune_s"], 3),
        "energy_scale_ms": (None if r["energy_scale_s"] is None
                            else round(1e3 * r["energy_scale_s"], 2)),
        "energy": r["energy"],
        # both D's of a step go to the host inside the timed region, as the
        # drop-in Engine's calls return them; the same step with D left on the
        # device would take 2 * d2h_ms_per_D less
        "pcie": {"d2h_ms_per_D": round(r["d2h_ms"], 3),
                 "D_bytes": int(H.n_blocks() * H.b_len ** 2 * 16),
                 "value_device_resident": N * M * S / (r["wall"] / args.steps
                                                      - 2e-3 * r["d2h_ms"])},
    }
    if not args.no_fp32 and args.precision == "fp64":
        _log("fp32-mixed run")
        r32 = run_gpu(H, groups, cfg, "fp32", args.steps, args.warmup, args.seed)
        t32 = r32["wall"] / args.steps
        ab32 = algorithmic_bytes(H, N, S, 8, table_K=(M + 1) // 2
                                 if r32["stoch_method"] == "table" else None)
        k32 = {}
        for k, (ms, n) in r32["kernel_ms"].items():
            if n:
                us = 1e3 * ms / n
                k32[k] = {"avg_launch_us": round(us, 3), "bytes_per_launch": int(ab32[k]),
                          "achieved_GBs": round(ab32[k] / (us * 1e-6) / 1e9, 1),
                          "frac": round(ab32[k] / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                          "ms_per_step": round(ms / args.steps, 3)}
        e_err = abs(r32["energy"] - r["energy"]) / abs(r["energy"])
        dg_err = float(np.max(np.abs(r32["Dg"] - r["Dg"])) / np.max(np.abs(r["Dg"])))
        ds_err = float(np.max(np.abs(r32["Ds"] - r["Ds"])) / np.max(np.abs(r["Ds"])))
        out["fp32_mixed"] = {"value": N * M * S / t32, "ms_per_step": round(1e3 * t32, 3),
                             "speedup_vs_fp64": round(r["wall"] / r32["wall"], 3),
                             "energy_rel_err_vs_fp64": e_err,
                             "dEdH_rel_err_vs_fp64": dg_err, "Dij_rel_err_vs_fp64": ds_err,
                             "kernels": k32, "kernel_config": r32["kernel_config"],
                             "stoch_method": r32["stoch_method"]}
        del r32
    if not args.no_serve:
        _log("serving stream")
        out["serving"] = serving_leg(args.serve_requests, 1 << 20, 2048, args.serve_rate,
                                     args.seed, args.serve_hbm_gb)
        sv = out["serving"]
        out["latency_ms"]["serving_fp64_p50"] = round(sv["fp64"]["p50_ms"], 3)
        out["latency_ms"]["serving_fp64_p99"] = round(sv["fp64"]["p99_ms"], 3)
        if "fp32" in sv:
            out["latency_ms"]["serving_fp32_p50"] = round(sv["fp32"]["p50_ms"], 3)
            out["latency_ms"]["serving_fp32_p99"] = round(sv["fp32"]["p99_ms"], 3)
    if not args.no_extras and args.config != "C2":
        _log("C2 leg")
        out["C2"] = extra_config_leg("C2", 3, 1, args.seed)
    if not args.no_extras:
        _log("C1 solve leg")
        out["C1"] = small_solve_leg("C1", args.seed)
    if not args.no_conductivity and args.precision == "fp64":
        # on the 2-D Kondo lattice (C2), whatever the headline workload
        _log(f"two-current moments, M={args.conductivity_moments}")
        Hc, gc, cfgc = build_workload("C2", args.seed)
        ec = HipEngine("cx_double", "fp64")
        esc = ec.energy_scale(Hc, 0.01)
        ec.close()
        out["conductivity"] = conductivity_leg(Hc, gc, cfgc, esc, args.seed,
                                               args.conductivity_moments)
        out["conductivity"]["workload"] = f"C2: {WORKLOADS['C2']}"
        del Hc
    _log("exact-diagonalisation check")
    out["exact_check"] = exact_check(args.config, args.precision, args.seed)
    if not args.no_cpu_baseline:
        m_cpu = args.cpu_sample_moments or max(2, 2 * round(2.5e8 / (N * S)))
        _log(f"cpu baseline (oracle, M'={m_cpu})")
        out["cpu_baseline"] = cpu_baseline(H, groups, cfg, r["es"], args.seed, m_cpu)
        threads = min(16, os.cpu_count() or 1, S)
        if threads > 1:
            _log(f"cpu baseline, {threads} column shards")
            out["cpu_baseline_threads"] = cpu_baseline_threads(
                H, groups, cfg, r["es"], args.seed, m_cpu, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
