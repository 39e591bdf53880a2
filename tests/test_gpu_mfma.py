"""Matrix-core (MFMA) kernels for complex 8x8 blocks, through the C-ABI,
against the CPU oracle on identical seeded inputs.  FP64: 1e-10 relative
(norm-wise, max|a-b|/max|b|); FP32-mixed: energy 1e-6, moments 2e-5 and D
1e-4 relative."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import fermi_coeffs, rel_err, seeded_R

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
K_MOM, K_XI, K_AD, K_OUT = 0, 1, 2, 3


def _dense_rows_b8(nb=24, extra=6, seed=3):
    """Hermitian 8x8-block matrix whose rows hold up to ~2*extra+1 blocks (more
    than one pass of the autodiff kernel's per-row block budget)."""
    from fastkpm_amd import lattice
    rng = np.random.default_rng(seed)
    rows, cols, blocks = [], [], []
    for i in range(nb):
        x = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
        rows.append(i)
        cols.append(i)
        blocks.append((x + x.conj().T) / 2)
        for j in rng.choice([k for k in range(nb) if k != i], size=extra, replace=False):
            t = 0.5 * (rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))
            rows += [i, int(j)]
            cols += [int(j), i]
            blocks += [t, t.conj().T]
    return lattice.coo_to_bsr(nb, 8, rows, cols, np.array(blocks))


def _lattice_b8(shape=(4, 4, 2), seed=7):
    from fastkpm_amd import lattice
    return lattice.lattice(shape, b=8, orbital_blocks=True, seed=seed)


def _setup(oracle, H, R, prec="fp64", mfma=True):
    from fastkpm_amd import HipEngine, EnergyScale
    lam = np.linalg.eigvalsh(H.to_dense())
    lo, hi = lam[0] - 0.1, lam[-1] + 0.1
    eng = HipEngine("cx_double", prec)
    eng.set_mfma(mfma)
    eng.set_R(R)
    eng.set_H(H, EnergyScale(lo, hi))
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    return eng, orc, lo, hi


def _probing_R(oracle, H, seed=11):
    groups = (np.arange(8)[None, :] + 8 * (np.arange(H.n_rows) % 3)[:, None]).reshape(-1)
    return seeded_R(oracle, groups.astype(np.int32), seed)


def _check_all(oracle, eng, orc, H, lo, hi, M, tol_mu, tol_d):
    c = fermi_coeffs(oracle, M, lo, hi, kT=0.05)
    assert rel_err(eng.moments(M), orc.moments(M)) < tol_mu
    D = H.like()
    eng.autodiff_matrix(c, D)
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < tol_d
    cf = fermi_coeffs(oracle, M, lo, hi, kT=0.05, kind=1)
    eng.stoch_matrix(cf, D)
    assert rel_err(D.val, orc.stoch_matrix(cf, D)) < tol_d


@pytest.mark.parametrize("case", ["lattice", "dense_rows"])
def test_mfma_fp64_matches_oracle(oracle, case):
    H = _lattice_b8() if case == "lattice" else _dense_rows_b8()
    assert case != "dense_rows" or int(np.max(np.diff(H.row_ptr))) > 8
    R = _probing_R(oracle, H)
    eng, orc, lo, hi = _setup(oracle, H, R)
    eng.moments(2)
    assert eng.kernel_config(K_MOM)["mfma"] and eng.kernel_config(K_AD)["mfma"]
    # rows of at most 8 blocks take the row-resident kernels, longer rows the general ones
    variant = 2 if case == "lattice" else 1
    assert eng.kernel_config(K_MOM)["mfma_variant"] == variant
    assert eng.kernel_config(K_AD)["mfma_variant"] == variant
    _check_all(oracle, eng, orc, H, lo, hi, 64, FP64_TOL, FP64_TOL)


@pytest.mark.parametrize("kind", ["dense_rows", "lattice"])
def test_mfma_equals_vector_kernels(oracle, kind):
    """The matrix-core (general and row-resident) and vector-unit paths agree
    with each other."""
    H = _dense_rows_b8(nb=16, extra=5, seed=9) if kind == "dense_rows" else _lattice_b8(seed=8)
    R = _probing_R(oracle, H, seed=5)
    out = {}
    for mf in (True, False):
        eng, orc, lo, hi = _setup(oracle, H, R, mfma=mf)
        M = 40
        c = fermi_coeffs(oracle, M, lo, hi)
        mu = eng.moments(M)
        assert eng.kernel_config(K_MOM)["mfma"] is mf
        D = H.like()
        eng.autodiff_matrix(c, D)
        out[mf] = (mu, D.val.copy())
    assert rel_err(out[True][0], out[False][0]) < 1e-12
    assert rel_err(out[True][1], out[False][1]) < 1e-12


@pytest.mark.parametrize("S_cols", [1, 3, 17, 40])
def test_mfma_ragged_columns(oracle, S_cols):
    """Dense R with a column count that leaves padded columns in the last tile."""
    H = _lattice_b8(shape=(3, 3, 2), seed=2)
    n = H.n_rows * H.b_len
    R = oracle.set_R_uncorrelated(n, S_cols, oracle.MT19937(4 + S_cols))
    eng, orc, lo, hi = _setup(oracle, H, R)
    assert eng.config()["s_padded"] % 16 == 0
    _check_all(oracle, eng, orc, H, lo, hi, 36, FP64_TOL, FP64_TOL)


@pytest.mark.parametrize("method", ["clenshaw", "forward"])
def test_mfma_stoch_methods_dense_R(oracle, method):
    H = _dense_rows_b8(nb=12, extra=4, seed=21)
    n = H.n_rows * H.b_len
    R = oracle.set_R_uncorrelated(n, 20, oracle.MT19937(8))
    eng, orc, lo, hi = _setup(oracle, H, R)
    eng.set_stoch_method(method)
    M = 50
    c = fermi_coeffs(oracle, M, lo, hi, kT=0.05, kind=1)
    D = H.like()
    eng.stoch_matrix(c, D)
    assert rel_err(D.val, orc.stoch_matrix(c, D)) < FP64_TOL
    for m in (2, 3):
        eng.stoch_matrix(c[:m], D)
        assert rel_err(D.val, orc.stoch_matrix(c[:m], D)) < FP64_TOL, m


def test_mfma_foreign_D_pattern(oracle):
    """D on the on-site blocks only: the unfused matrix-core autodiff step plus
    the sparse outer product."""
    from fastkpm_amd import lattice
    H = _lattice_b8(shape=(3, 3, 2), seed=12)
    nb = H.n_rows
    D = lattice.coo_to_bsr(nb, 8, list(range(nb)), list(range(nb)), np.zeros((nb, 8, 8)))
    R = _probing_R(oracle, H, seed=3)
    eng, orc, lo, hi = _setup(oracle, H, R)
    M = 44
    eng.moments(M)
    orc.moments(M)
    c = fermi_coeffs(oracle, M, lo, hi)
    eng.autodiff_matrix(c, D)
    assert eng.config()["fused_autodiff"] is False
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < FP64_TOL


def test_mfma_fp32_mixed(oracle):
    """FP32 matrix-core recursion with FP64 dots and D against the FP64 oracle."""
    H = _lattice_b8(shape=(4, 4, 2), seed=31)
    R = _probing_R(oracle, H, seed=6)
    eng, orc, lo, hi = _setup(oracle, H, R, prec="fp32")
    M = 80
    c = fermi_coeffs(oracle, M, lo, hi, kT=0.05)
    mu = eng.moments(M)
    mu_o = orc.moments(M)
    assert eng.kernel_config(K_MOM)["mfma_variant"] == 2
    assert rel_err(mu, mu_o) < 2e-5
    e, e_o = float(np.dot(c, mu)), float(np.dot(c, mu_o))
    assert abs(e - e_o) / abs(e_o) < 1e-6
    D = H.like()
    eng.autodiff_matrix(c, D)
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < 1e-4
    eng.stoch_matrix(c, D)
    assert rel_err(D.val, orc.stoch_matrix(c, D)) < 1e-4


@pytest.mark.parametrize("S_cols,slots", [(8, 7), (24, 7), (24, 8), (40, 8)])
def test_mfma_fp32_row_resident_autodiff(oracle, S_cols, slots):
    """The FP32 row-resident autodiff step (FKPM_OPT_MFMA 2) on single
    16-column tiles (S = 8: a padded width of 8 packets) and on pairs of tiles
    (S = 24, 40: 16 and 32 packets), with 7 and 8 block slots per row,
    against the FP64 oracle."""
    from fastkpm_amd import lattice
    from helpers import with_partner_bonds
    shape = (4, 4, 4)
    H = lattice.lattice(shape, b=8, orbital_blocks=True, seed=5)
    if slots == 8:
        H = with_partner_bonds(H, shape, 6)
    assert int(np.max(np.diff(H.row_ptr))) == slots
    n = H.n_rows * H.b_len
    R = oracle.set_R_uncorrelated(n, S_cols, oracle.MT19937(3 + S_cols))
    eng, orc, lo, hi = _setup(oracle, H, R, prec="fp32", mfma=2)
    eng.moments(2)
    assert eng.kernel_config(K_AD)["mfma_variant"] == 2
    packets = eng.config()["s_padded"] // 2  # two complex columns per 16-byte packet
    assert (packets % 16 == 0) == (S_cols > 16)
    _check_all(oracle, eng, orc, H, lo, hi, 40, 2e-5, 1e-4)
    eng.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mfma_forced_row_resident_falls_back_on_long_rows(oracle, prec):
    """FKPM_OPT_MFMA 2 asks for the row-resident kernels; rows of more than 8
    blocks have none, so every kind keeps the general matrix-core kernels and
    the results stay those of the oracle."""
    H = _dense_rows_b8(nb=16, extra=5, seed=4)
    assert int(np.max(np.diff(H.row_ptr))) > 8
    R = _probing_R(oracle, H, seed=9)
    eng, orc, lo, hi = _setup(oracle, H, R, prec=prec, mfma=2)
    eng.moments(2)
    assert eng.kernel_config(K_MOM)["mfma_variant"] == 1
    assert eng.kernel_config(K_AD)["mfma_variant"] == 1
    tol = (FP64_TOL, FP64_TOL) if prec == "fp64" else (2e-5, 1e-4)
    _check_all(oracle, eng, orc, H, lo, hi, 32, *tol)
    eng.close()


@pytest.mark.parametrize("shape", [(64, 64), (16, 16, 16)])
def test_mfma_fp32_pairs_of_tiles_at_4096_rows(shape):
    """FP32 row-resident forward and autodiff steps on pairs of column tiles
    (S = 32 and 64 probing colours) on 4096 block rows of a 2-D (5 blocks a
    row: two padding slots) and a 3-D lattice, against the FP64 engine:
    moments within 2e-5, energy within 1e-6, dE/dH within 1e-4."""
    from fastkpm_amd import HipEngine, host, lattice
    from fastkpm_amd.serving import gershgorin
    H = lattice.lattice(shape, b=8, orbital_blocks=True, seed=12)
    assert H.n_rows == 4096
    groups = lattice.colouring(shape, (2,) * len(shape), 8)
    R = host.set_R_correlated("cx_double", groups, host.RNG(12))
    es = gershgorin(H)
    M = 64
    c = host.expansion_coefficients(M, 2 * M, host.FERMI_ENERGY, 0.05, 0.0, es.lo, es.hi)
    out = {}
    for prec in ("fp64", "fp32"):
        eng = HipEngine("cx_double", prec)
        eng.set_R(R)
        eng.set_H(H, es)
        mu = eng.moments(M)
        assert eng.kernel_config(K_MOM)["mfma_variant"] == 2
        assert eng.kernel_config(K_AD)["mfma_variant"] == 2
        D = H.like()
        eng.autodiff_matrix(c, D)
        out[prec] = (mu, float(np.dot(c, mu)), D.val.copy())
        if prec == "fp32":  # two complex columns per packet, 16 packets per pair of tiles
            assert eng.config()["s_padded"] // 2 % 16 == 0
        eng.close()
    assert rel_err(out["fp32"][0], out["fp64"][0]) < 2e-5
    assert abs(out["fp32"][1] - out["fp64"][1]) < 1e-6 * abs(out["fp64"][1])
    assert rel_err(out["fp32"][2], out["fp64"][2]) < 1e-4


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mfma_autotuned(oracle, prec):
    H = _lattice_b8(shape=(4, 4, 4), seed=5)
    R = _probing_R(oracle, H, seed=2)
    eng, orc, lo, hi = _setup(oracle, H, R, prec=prec)
    eng.autotune()
    assert eng.state()["tuned"]
    if prec == "fp64":
        _check_all(oracle, eng, orc, H, lo, hi, 32, FP64_TOL, FP64_TOL)
    else:
        _check_all(oracle, eng, orc, H, lo, hi, 32, 2e-5, 1e-4)
