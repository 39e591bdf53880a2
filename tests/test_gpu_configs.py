"""The measured configurations and all four engine element types on the HIP
path, against the oracle on identical inputs.

- C1 at its full size: 32 x 32 periodic square lattice, t = -1, real H,
  energy scale [-4.1, 4.1], M = 500, Fermi functions at mu = 0 and kT = 0 (the
  step / kink branch of fermi_energy and fermi_density, src/fastkpm.cpp:375-393),
  4 x 4 probing colours (S = 16), through a `double` engine (real +-1 phases)
  and a `cx_double` engine (Z4 phases): moments, energy, autodiff dE/dH and
  stoch_matrix D on the nearest-neighbour pattern at 1e-10; with R = I the
  moments equal the closed form sum_k T_m(eps_k / 4.1), eps_k = -2 (cos kx + cos ky).
- Engine<float> and Engine<cx_float> (src/engine.cpp:118-121) through the C-ABI
  (FKPM_F32 / FKPM_C32) in both storage precisions.
- A second set_H on the same handle with longer block rows (b = 8: 7 blocks
  per row, then 11) picks the kernels again.

Tolerances are norm-wise (max |a - b| / max |b|).  FP64 storage: 1e-10.
Single-precision results: D values returned as float are rounded once
(2^-24 relative); FP32 storage of the recursion accumulates at most about
16 roundings of 2^-24 per element and step over M/2 steps, so moments and D
are held to M * 2^-24 * 16 and the energy to 1e-6.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import rel_err, seeded_R, with_partner_bonds

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
U32 = 2.0 ** -24


def _fp32_tol(M: int) -> float:
    return M * U32 * 16


def _c1(dtype):
    from fastkpm_amd import lattice
    H, groups, cfg = lattice.make_config("C1", seed=0, dtype=dtype)
    return H, groups, cfg


def _coeffs(M, lo, hi, kT, mu):
    from fastkpm_amd import host
    cE = host.expansion_coefficients(M, 2 * M, host.FERMI_ENERGY, kT, mu, lo, hi)
    cF = host.expansion_coefficients(M, 2 * M, host.FERMI_DENSITY, kT, mu, lo, hi)
    return cE, cF


@pytest.mark.parametrize("etype", ["double", "cx_double"])
def test_c1_full_size_matches_oracle(oracle, etype):
    from fastkpm_amd import EnergyScale, HipEngine, host
    real = etype == "double"
    H, groups, cfg = _c1(np.float64 if real else np.complex128)
    M, (lo, hi), kT, mu = cfg["M"], cfg["es"], cfg["kT"], cfg["mu"]
    assert (H.n_rows, H.b_len, M, kT) == (1024, 1, 500, 0.0)
    eng = HipEngine(etype, "fp64")
    eng.set_R_correlated(groups, host.RNG(2024))
    assert eng.R.shape == (1024, 16)
    eng.set_H(H, EnergyScale(lo, hi))
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(eng.R)
    # the product's R is the oracle's restatement of set_R_correlated
    R_o = oracle.set_R_correlated(groups, oracle.MT19937(2024), is_complex=not real)
    assert np.array_equal(eng.R, R_o.real if real else R_o)
    cE, cF = _coeffs(M, lo, hi, kT, mu)
    # kT = 0: c_m of the kink x * theta(-x); its Chebyshev series at x = 0 is
    # what the step branch of fermi_energy gives
    assert np.all(np.isfinite(cE)) and np.all(np.isfinite(cF))
    mu_g = eng.moments(M)
    mu_o = orc.moments(M)
    assert rel_err(mu_g, mu_o) < FP64_TOL
    assert mu_g[0] == pytest.approx(1024.0, rel=1e-14)
    e_g, e_o = float(np.dot(cE, mu_g)), float(np.dot(cE, mu_o))
    assert abs(e_g - e_o) < FP64_TOL * abs(e_o)
    D = H.like()
    eng.autodiff_matrix(cE, D)
    assert rel_err(D.val, orc.autodiff_matrix(cE, D)) < FP64_TOL
    D2 = H.like()
    eng.stoch_matrix(cF, D2)
    assert rel_err(D2.val, orc.stoch_matrix(cF, D2)) < FP64_TOL
    # the diagonal of D = f(H) sums to the stochastic particle count
    # sum_m cF_m mu_m (mu = 0 at half filling: about N / 2)
    diag = np.nonzero(D2.row_idx == D2.col_idx)[0]
    n_part = float(np.sum(D2.val[diag].real))
    assert n_part == pytest.approx(float(np.dot(cF, mu_g)), rel=1e-10)
    assert 400 < n_part < 624


def test_c1_identity_R_matches_closed_form():
    """R = I (S = N = 1024 columns): mu_m = tr T_m(Hs) = sum_k T_m(eps_k / 4.1)
    over the 32 x 32 Brillouin zone, eps_k = -2 (cos kx + cos ky)."""
    from fastkpm_amd import EnergyScale, HipEngine
    H, _, cfg = _c1(np.float64)
    M, (lo, hi) = cfg["M"], cfg["es"]
    eng = HipEngine("double", "fp64")
    eng.set_R(np.eye(1024))
    eng.set_H(H, EnergyScale(lo, hi))
    mu_g = eng.moments(M)
    k = 2 * np.pi * np.arange(32) / 32
    eps = (-2.0 * (np.cos(k)[:, None] + np.cos(k)[None, :])).reshape(-1)
    th = np.arccos(eps / hi)  # avg = 0, mag = 4.1
    mu_c = np.cos(np.arange(M)[:, None] * th[None, :]).sum(axis=1)
    assert rel_err(mu_g, mu_c) < FP64_TOL


SMALL = dict(shape=(8, 8), periods=(4, 4))


def _typed_case(oracle, etype):
    """An 8 x 8 lattice in the engine's element type: b = 1 real with on-site
    disorder for float, b = 2 Kondo for cx_float."""
    from fastkpm_amd import lattice
    real = etype in ("float", "double")
    if real:
        H = lattice.lattice(SMALL["shape"], b=1, W=2.0, seed=4, dtype=np.float64)
        groups = lattice.colouring(SMALL["shape"], SMALL["periods"], 1)
    else:
        H = lattice.lattice(SMALL["shape"], b=2, J=2.0, seed=4)
        groups = lattice.colouring(SMALL["shape"], SMALL["periods"], 2)
    np_t = np.float32 if etype == "float" else np.complex64
    H.val = H.val.astype(np_t)
    R = seeded_R(oracle, groups, 5, is_complex=not real)
    R = (R.real if real else R).astype(np_t)
    return H, R


@pytest.mark.parametrize("etype", ["float", "cx_float"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_single_precision_engine_types(oracle, etype, prec):
    """Engine<float> / Engine<cx_float>: host data in single precision through
    FKPM_F32 / FKPM_C32.  The oracle runs in FP64 on the same (exactly widened)
    inputs.  fp64 storage: moments at 1e-10, D rounded once to float; fp32
    storage: the FP32 recursion bound."""
    from fastkpm_amd import EnergyScale, HipEngine
    H, R = _typed_case(oracle, etype)
    lam = np.linalg.eigvalsh(H.to_dense().astype(np.complex128))
    lo, hi = float(lam[0]) - 0.2, float(lam[-1]) + 0.2
    M = 128
    eng = HipEngine(etype, prec)
    eng.set_R(R)
    eng.set_H(H, EnergyScale(lo, hi))
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    cE, cF = _coeffs(M, lo, hi, 0.05, 0.0)
    tol_mu = FP64_TOL if prec == "fp64" else _fp32_tol(M)
    tol_d = 4 * U32 if prec == "fp64" else _fp32_tol(M)
    mu_g = eng.moments(M)
    mu_o = orc.moments(M)
    assert rel_err(mu_g, mu_o) < tol_mu
    e_g, e_o = float(np.dot(cE, mu_g)), float(np.dot(cE, mu_o))
    assert abs(e_g - e_o) < (FP64_TOL if prec == "fp64" else 1e-6) * abs(e_o)
    D = H.like()
    eng.autodiff_matrix(cE, D)
    assert D.val.dtype == H.val.dtype
    assert rel_err(D.val, orc.autodiff_matrix(cE, D)) < tol_d
    D2 = H.like()
    eng.stoch_matrix(cF, D2)
    assert rel_err(D2.val, orc.stoch_matrix(cF, D2)) < tol_d


def _with_second_neighbours(H1, shape, seed):
    """H1 plus Hermitian pairs of random 8 x 8 blocks on the (+1, +1, 0) and
    (+1, -1, 0) bonds: 7 + 4 = 11 blocks per row."""
    from fastkpm_amd import lattice
    rng = np.random.default_rng(seed)
    b = H1.b_len
    nsite = H1.n_rows
    idx = np.arange(nsite).reshape(shape, order="F")
    rows, cols, blks = [H1.row_idx], [H1.col_idx], [H1.val.reshape(-1, b, b).transpose(0, 2, 1)]
    for sy in (1, -1):
        dst = np.roll(np.roll(idx, -1, axis=0), -sy, axis=1).reshape(-1, order="F")
        src = idx.reshape(-1, order="F")
        T = (rng.standard_normal((nsite, b, b)) + 1j * rng.standard_normal((nsite, b, b))) * 0.3
        rows += [src, dst]
        cols += [dst, src]
        blks += [T, np.conj(T.transpose(0, 2, 1))]
    return lattice.coo_to_bsr(nsite, b, np.concatenate(rows), np.concatenate(cols),
                              np.concatenate(blks, axis=0))


def test_set_H_with_longer_rows_picks_kernels_again(oracle):
    """b = 8 complex: the row-resident matrix-core kernels serve rows of at most
    8 blocks.  A second set_H on the same handle with 11 blocks per row must
    leave them (same b, same block rows, same R)."""
    from fastkpm_amd import EnergyScale, HipEngine, lattice
    shape = (4, 4, 4)
    H1 = lattice.lattice(shape, b=8, orbital_blocks=True, seed=9)
    H2 = _with_second_neighbours(H1, shape, 10)
    deg = lambda H: int(np.max(np.diff(H.row_ptr)))
    assert deg(H1) == 7 and deg(H2) == 11 and H2.n_rows == H1.n_rows
    groups = lattice.colouring(shape, (2, 2, 2), 8)
    R = seeded_R(oracle, groups, 3)
    M = 48
    eng = HipEngine("cx_double", "fp64")
    eng.set_R(R)
    orc = oracle.Engine()
    orc.set_R(R)
    for H in (H1, H2, H1):
        lam = np.linalg.eigvalsh(H.to_dense())
        lo, hi = float(lam[0]) - 0.5, float(lam[-1]) + 0.5
        eng.set_H(H, EnergyScale(lo, hi))
        orc.set_H(H, lo, hi)
        mu_g = eng.moments(M)
        assert rel_err(mu_g, orc.moments(M)) < FP64_TOL, deg(H)
        if deg(H) > 8:
            assert eng.kernel_config(0)["mfma_variant"] != 2
        cE, cF = _coeffs(M, lo, hi, 0.1, 0.0)
        D = H.like()
        eng.autodiff_matrix(cE, D)
        assert rel_err(D.val, orc.autodiff_matrix(cE, D)) < FP64_TOL, deg(H)
        D2 = H.like()
        eng.stoch_matrix(cF, D2)
        assert rel_err(D2.val, orc.stoch_matrix(cF, D2)) < FP64_TOL, deg(H)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_row_resident_seven_and_eight_block_rows(oracle, prec):
    """The row-resident matrix-core kernels come in two instantiations: 7 block
    slots per row (a cubic lattice: no padding slot) and 8.  One handle, set_H
    with 7, then 8, then 7 blocks per row: each pass against the oracle."""
    from fastkpm_amd import EnergyScale, HipEngine, lattice
    shape = (4, 4, 4)
    H7 = lattice.lattice(shape, b=8, orbital_blocks=True, seed=5)
    H8 = with_partner_bonds(H7, shape, 6)
    deg = lambda H: int(np.max(np.diff(H.row_ptr)))
    assert deg(H7) == 7 and deg(H8) == 8 and int(np.min(np.diff(H8.row_ptr))) == 8
    groups = lattice.colouring(shape, (2, 2, 2), 8)
    R = seeded_R(oracle, groups, 4)
    M = 40
    tol_mu, tol_d = (FP64_TOL, FP64_TOL) if prec == "fp64" else (2e-5, 1e-4)
    eng = HipEngine("cx_double", prec)
    eng.set_R(R)
    orc = oracle.Engine()
    orc.set_R(R)
    for H in (H7, H8, H7):
        lam = np.linalg.eigvalsh(H.to_dense())
        lo, hi = float(lam[0]) - 0.5, float(lam[-1]) + 0.5
        eng.set_H(H, EnergyScale(lo, hi))
        orc.set_H(H, lo, hi)
        assert rel_err(eng.moments(M), orc.moments(M)) < tol_mu, deg(H)
        assert eng.kernel_config(0)["mfma_variant"] == 2
        cE, cF = _coeffs(M, lo, hi, 0.1, 0.0)
        D = H.like()
        eng.autodiff_matrix(cE, D)
        assert rel_err(D.val, orc.autodiff_matrix(cE, D)) < tol_d, deg(H)
        D2 = H.like()
        eng.stoch_matrix(cF, D2)
        assert rel_err(D2.val, orc.stoch_matrix(cF, D2)) < tol_d, deg(H)
    eng.close()


@pytest.mark.parametrize("method,store", [("auto", 2), ("clenshaw", 0), ("forward", 0)])
def test_step_loops_as_graphs_equal_direct_launches(oracle, method, store):
    """FKPM_OPT_GRAPH: the C1 solve (moments, autodiff_matrix, stoch_matrix)
    with every step loop replayed as a cached hipGraph is bitwise the solve
    with direct launches, on every replay, with other eager work (a different
    M, so other graphs and buffers) between the replays; the first solve is
    also checked against the oracle."""
    from fastkpm_amd import EnergyScale, HipEngine, host
    H, groups, cfg = _c1(np.complex128)
    M, (lo, hi), kT, mu = cfg["M"], cfg["es"], cfg["kT"], cfg["mu"]
    cE, cF = _coeffs(M, lo, hi, kT, mu)
    c2E, _ = _coeffs(64, lo, hi, kT, mu)

    def solve(eng, m=M, ce=cE, cf=cF):
        mu_ = eng.moments(m)
        D1, D2 = H.like(), H.like()
        eng.autodiff_matrix(ce, D1)
        eng.stoch_matrix(cf, D2)
        return mu_, D1.val, D2.val

    engines = {}
    for g in (0, 1):
        eng = HipEngine("cx_double", "fp64")
        eng.set_graph(g)
        eng.set_store_alpha(store)
        eng.set_stoch_method(method)
        eng.set_R_correlated(groups, host.RNG(2024))
        eng.set_H(H, EnergyScale(lo, hi))
        engines[g] = eng
    ref = solve(engines[0])
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(engines[0].R)
    assert rel_err(ref[0], orc.moments(M)) < FP64_TOL
    for rep in range(4):
        out = solve(engines[1])
        for a, b in zip(out, ref):
            assert np.array_equal(a, b), rep
        solve(engines[1], 64, c2E, c2E)  # eager work on other graphs in between
        engines[0].synchronize()
    st = engines[1].graph_stats()
    # three step loops per solve, two solve shapes: each loop captured once,
    # then replayed (4 repetitions x 2 solves x 3 loops launches)
    assert st["graph_captures"] == 6 and st["graph_launches"] == 24
    assert engines[0].graph_stats()["graph_launches"] == 0
