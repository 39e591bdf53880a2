"""Two-current moments (Engine::moments2_v1) on the device through the C-ABI,
against the CPU oracle on identical seeded inputs: FP64 1e-10 relative
(norm-wise, max|a-b|/max|b|), FP32-mixed 2e-5; probe-column chunking (by the
caller's chunk and by the memory budget) changes nothing beyond rounding."""
from __future__ import annotations

import os

import numpy as np
import pytest

from helpers import rel_err, seeded_R

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _case(oracle, kind, seed=3):
    """(H, lo, hi, R, j1, j2, dtype)."""
    from fastkpm_amd import lattice
    if kind == "b2":
        shape, per = (6, 8), (3, 4)
        H = lattice.lattice(shape, b=2, J=1.5, W=1.0, seed=seed)
        groups = lattice.colouring(shape, per, 2)
        R = seeded_R(oracle, groups, seed + 10)
        j1, j2 = lattice.current(H, shape, 0), lattice.current(H, shape, 1)
        dt = "cx_double"
    elif kind == "b8":
        shape, per = (3, 3, 4), (3, 3, 2)
        H = lattice.lattice(shape, b=8, orbital_blocks=True, seed=seed)
        groups = lattice.colouring(shape, per, 8)
        R = seeded_R(oracle, groups, seed + 10)
        j1 = j2 = lattice.current(H, shape, 2)
        dt = "cx_double"
    elif kind == "b1_real":
        shape, per = (8, 6), (4, 3)
        H = lattice.lattice(shape, b=1, W=2.0, seed=seed, dtype=np.float64)
        groups = lattice.colouring(shape, per, 1)
        R = seeded_R(oracle, groups, seed + 10, is_complex=False).real
        j1, j2 = lattice.current(H, shape, 0, real=True), lattice.current(H, shape, 1, real=True)
        dt = "double"
    else:  # dense R, b = 4, a current with its own pattern and the identity
        shape = (4, 3)
        H = lattice.lattice(shape, b=4, orbital_blocks=True, seed=seed)
        n = H.n_rows * 4
        R = oracle.set_R_uncorrelated(n, 7, oracle.MT19937(seed))
        j1 = lattice.current(H, shape, 0)
        j2 = lattice.coo_to_bsr(H.n_rows, 4, list(range(H.n_rows)), list(range(H.n_rows)),
                                np.array([np.eye(4)] * H.n_rows))
        dt = "cx_double"
    lam = np.linalg.eigvalsh(H.to_dense())
    return H, lam[0] - 0.1, lam[-1] + 0.1, R, j1, j2, dt


def _engines(oracle, H, lo, hi, R, dt, prec="fp64"):
    from fastkpm_amd import HipEngine, EnergyScale
    eng = HipEngine(dt, prec)
    eng.set_R(R)
    eng.set_H(H, EnergyScale(lo, hi))
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    return eng, orc


@pytest.mark.parametrize("kind", ["b2", "b8", "b1_real", "b4_dense_R"])
def test_moments2_fp64_matches_oracle(oracle, kind):
    H, lo, hi, R, j1, j2, dt = _case(oracle, kind)
    eng, orc = _engines(oracle, H, lo, hi, R, dt)
    M = 24
    mu2 = eng.moments2_v1(M, j1, j2)
    mu2_o = orc.moments2_v1(M, j1, j2)
    print(kind, float(np.max(np.abs(mu2_o))), rel_err(mu2, mu2_o), flush=True)
    assert np.max(np.abs(mu2_o)) > 0
    assert rel_err(mu2, mu2_o) < FP64_TOL
    if dt == "double":
        assert not np.any(mu2.imag)


@pytest.mark.parametrize("chunk", [1, 3, 5])
def test_moments2_column_chunks(oracle, chunk):
    H, lo, hi, R, j1, j2, dt = _case(oracle, "b2", seed=5)
    eng, orc = _engines(oracle, H, lo, hi, R, dt)
    M = 16
    full = eng.moments2_v1(M, j1, j2)
    part = eng.moments2_v1(M, j1, j2, col_chunk=chunk)
    assert rel_err(part, full) < 1e-13
    assert rel_err(part, orc.moments2_v1(M, j1, j2)) < FP64_TOL


def test_moments2_budget_splits_and_refuses(oracle):
    """A memory budget below one pass over all columns splits the columns;
    one below a single column is refused."""
    from fastkpm_amd._capi import FkpmError
    H, lo, hi, R, j1, j2, dt = _case(oracle, "b2", seed=6)
    eng, orc = _engines(oracle, H, lo, hi, R, dt)
    M = 12
    full = eng.moments2_v1(M, j1, j2)
    n, S = H.n_rows * 2, R.shape[1]
    assert S == 24  # one pass pads to 32 columns, half of them to 16
    per_col = (2 * M + 3) * n * 16  # (2M+3) vectors of one complex-double column
    os.environ["FKPM_MOMENTS2_BYTES"] = str(per_col * S)
    try:
        split = eng.moments2_v1(M, j1, j2)
        # a single column still pads to the narrowest pass width (4 columns)
        os.environ["FKPM_MOMENTS2_BYTES"] = str(per_col * 2)
        with pytest.raises(FkpmError, match="budget"):
            eng.moments2_v1(M, j1, j2)
    finally:
        del os.environ["FKPM_MOMENTS2_BYTES"]
    assert rel_err(split, full) < 1e-13


def test_moments2_fp32_mixed(oracle):
    H, lo, hi, R, j1, j2, dt = _case(oracle, "b2", seed=7)
    eng, orc = _engines(oracle, H, lo, hi, R, dt, prec="fp32")
    M = 32
    assert rel_err(eng.moments2_v1(M, j1, j2), orc.moments2_v1(M, j1, j2)) < 2e-5


def test_moments2_keeps_moments_state(oracle):
    """moments(M), moments2_v1, then autodiff_matrix with the same M."""
    from helpers import fermi_coeffs
    H, lo, hi, R, j1, j2, dt = _case(oracle, "b2", seed=8)
    eng, orc = _engines(oracle, H, lo, hi, R, dt)
    M = 20
    eng.moments(M)
    orc.moments(M)
    eng.moments2_v1(8, j1, j2)
    c = fermi_coeffs(oracle, M, lo, hi, kT=0.05)
    D = H.like()
    eng.autodiff_matrix(c, D)
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < FP64_TOL


def test_moments2_identity_currents(oracle):
    """j1 = j2 = I: mu2[a][b] = (mu_{a+b} + mu_{|a-b|}) / 2 with the device's own moments."""
    from fastkpm_amd import lattice
    H, lo, hi, R, _, _, dt = _case(oracle, "b2", seed=9)
    eng, _ = _engines(oracle, H, lo, hi, R, dt)
    I = lattice.coo_to_bsr(H.n_rows, 2, list(range(H.n_rows)), list(range(H.n_rows)),
                           np.array([np.eye(2)] * H.n_rows))
    M = 14
    mu2 = eng.moments2_v1(M, I, I)
    mu = eng.moments(2 * M)
    ref = np.array([[0.5 * (mu[a + b] + mu[abs(a - b)]) for b in range(M)] for a in range(M)])
    assert rel_err(mu2, ref) < 1e-12


def test_moments2_static_conductivity(oracle):
    """sigma_xx = sum c[m1][m2] mu2[m1][m2] from the device moments and the
    host coefficients equals the same sum over the oracle's moments."""
    from fastkpm_amd import host
    H, lo, hi, R, j1, _, dt = _case(oracle, "b2", seed=10)
    eng, orc = _engines(oracle, H, lo, hi, R, dt)
    M = 32
    c2 = host.conductivity_coefficients(M, 2 * M, 0.1, 0.0, 0.0, lo, hi)
    s = np.sum(c2 * eng.moments2_v1(M, j1, j1))
    s_o = np.sum(c2 * orc.moments2_v1(M, j1, j1))
    assert abs(s - s_o) < 1e-10 * abs(s_o)
    assert s_o.real > 0


def test_moments2_refusals(oracle):
    from fastkpm_amd import lattice
    from fastkpm_amd._capi import FkpmError
    H, lo, hi, R, j1, j2, dt = _case(oracle, "b2", seed=11)
    eng, _ = _engines(oracle, H, lo, hi, R, dt)
    with pytest.raises(FkpmError, match="even"):
        eng.moments2_v1(7, j1, j2)
    bad = lattice.coo_to_bsr(H.n_rows, 2, [0], [H.n_rows - 1], np.ones((1, 2, 2)))
    bad.col_idx = bad.col_idx.copy()
    bad.col_idx[0] = H.n_rows  # out of range
    with pytest.raises(FkpmError):
        eng.moments2_v1(8, bad, j2)
