"""Parity of the HIP engine (through the C-ABI) with the CPU oracle on
identical seeded inputs.  FP64: 1e-10 relative (norm-wise, max|a-b|/max|b|).
FP32-mixed: the tolerance stated in each test."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import fermi_coeffs, rel_err, seeded_R

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _engine(dtype="cx_double", prec="fp64"):
    from fastkpm_amd import HipEngine
    return HipEngine(dtype, prec)


def _case(oracle, shape, b, seed, J=0.0, W=0.0, orbital_blocks=False, periods=None,
          real=False):
    from fastkpm_amd import lattice
    dtype = np.float64 if real else np.complex128
    H = lattice.lattice(shape, b=b, J=J, W=W, orbital_blocks=orbital_blocks, seed=seed,
                        dtype=dtype)
    periods = periods or tuple(2 for _ in shape)
    groups = lattice.colouring(shape, periods, b)
    R = seeded_R(oracle, groups, seed + 11, is_complex=not real)
    if real:
        R = R.real
    return H, R


def _run_pair(oracle, H, R, lo, hi, M, dtype="cx_double", prec="fp64", D=None):
    eng = _engine(dtype, prec)
    es_lo, es_hi = lo, hi
    from fastkpm_amd import EnergyScale
    eng.set_R(R)
    eng.set_H(H, EnergyScale(es_lo, es_hi))
    orc = oracle.Engine()
    orc.set_H(H, es_lo, es_hi)
    orc.set_R(R)
    D = D if D is not None else H.like()
    return eng, orc, D


CASES = [
    # name, shape, b, J, W, orbital_blocks, periods, real
    ("square_b1_real", (8, 8), 1, 0.0, 1.0, False, (4, 4), True),
    ("square_b1_cx", (8, 8), 1, 0.0, 2.0, False, (4, 4), False),
    ("kondo_b2", (8, 8), 2, 2.0, 0.0, False, (4, 4), False),
    ("anderson_kondo_3d_b2", (4, 4, 4), 2, 1.0, 4.0, False, (2, 2, 2), False),
    ("orbital_b4", (4, 4), 4, 0.0, 0.0, True, (2, 2), False),
    ("orbital_b8", (4, 4, 2), 8, 0.0, 0.0, True, (2, 2, 2), False),
]


@pytest.mark.parametrize("name,shape,b,J,W,ob,periods,real", CASES, ids=[c[0] for c in CASES])
def test_fp64_moments_stoch_autodiff(oracle, name, shape, b, J, W, ob, periods, real):
    H, R = _case(oracle, shape, b, seed=7, J=J, W=W, orbital_blocks=ob, periods=periods, real=real)
    lam = np.linalg.eigvalsh(H.to_dense())
    lo, hi = lam[0] - 0.1, lam[-1] + 0.1
    M = 64
    dtype = "double" if real else "cx_double"
    eng, orc, D = _run_pair(oracle, H, R, lo, hi, M, dtype=dtype)
    c = fermi_coeffs(oracle, M, lo, hi, kT=0.05)

    mu_g = eng.moments(M)
    mu_o = orc.moments(M)
    assert rel_err(mu_g, mu_o) < FP64_TOL, (name, rel_err(mu_g, mu_o))

    eng.autodiff_matrix(c, D)
    d_o = orc.autodiff_matrix(c, D)
    assert rel_err(D.val, d_o) < FP64_TOL, (name, "autodiff", rel_err(D.val, d_o))

    D2 = H.like()
    eng.stoch_matrix(c, D2)
    s_o = orc.stoch_matrix(c, D2)
    assert rel_err(D2.val, s_o) < FP64_TOL, (name, "stoch", rel_err(D2.val, s_o))


@pytest.mark.parametrize("S_cols", [1, 3, 5, 16, 33])
def test_fp64_ragged_columns(oracle, S_cols):
    """R with a column count that is not a multiple of any lane geometry."""
    from fastkpm_amd import lattice, EnergyScale
    H = lattice.lattice((6, 6), b=2, J=1.5, seed=3)
    n = H.n_rows * H.b_len
    rng = oracle.MT19937(99)
    R = oracle.set_R_uncorrelated(n, S_cols, rng)
    lo, hi = -6.0, 6.0
    eng = _engine()
    eng.set_R(R)
    eng.set_H(H, EnergyScale(lo, hi))
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    M = 40
    assert rel_err(eng.moments(M), orc.moments(M)) < FP64_TOL
    c = fermi_coeffs(oracle, M, lo, hi)
    D = H.like()
    eng.autodiff_matrix(c, D)
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < FP64_TOL


def test_fp64_foreign_D_pattern(oracle):
    """D on a pattern other than H's (second neighbours): the unfused path."""
    from fastkpm_amd import lattice, EnergyScale
    shape = (6, 6)
    H = lattice.lattice(shape, b=2, J=2.0, seed=5)
    # pattern: on-site + next-nearest (diagonal) neighbours, symmetric
    P = lattice.lattice(shape, b=2, t=0.0, seed=5)
    nsite = 36
    rows, cols = [], []
    for s in range(nsite):
        x, y = s % 6, s // 6
        for dx, dy in ((0, 0), (1, 1), (-1, -1), (1, -1), (-1, 1)):
            rows.append(s)
            cols.append(((x + dx) % 6) + 6 * ((y + dy) % 6))
    D = lattice.coo_to_bsr(nsite, 2, rows, cols, np.zeros((len(rows), 2, 2)))
    del P
    groups = lattice.colouring(shape, (3, 3), 2)
    R = seeded_R(oracle, groups, 4)
    lo, hi = -7.0, 7.0
    eng = _engine()
    eng.set_R(R)
    eng.set_H(H, EnergyScale(lo, hi))
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    M = 50
    eng.moments(M)
    orc.moments(M)
    c = fermi_coeffs(oracle, M, lo, hi)
    eng.autodiff_matrix(c, D)
    assert eng.config()["fused_autodiff"] is False
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < FP64_TOL
    eng.stoch_matrix(c, D)
    assert rel_err(D.val, orc.stoch_matrix(c, D)) < FP64_TOL


def test_fp32_mixed_energy(oracle):
    """FP32 recursion storage, FP64 accumulation: energy within 1e-6 relative."""
    from fastkpm_amd import lattice, EnergyScale
    H = lattice.lattice((16, 16), b=2, J=2.0, seed=1)
    groups = lattice.colouring((16, 16), (4, 8), 2)
    R = seeded_R(oracle, groups, 21)
    lo, hi = -6.5, 6.5
    M = 400
    eng = _engine("cx_double", "fp32")
    eng.set_R(R)
    eng.set_H(H, EnergyScale(lo, hi))
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    mu_g = eng.moments(M)
    mu_o = orc.moments(M)
    c = fermi_coeffs(oracle, M, lo, hi)
    e_g = float(np.dot(c, mu_g))
    e_o = float(np.dot(c, mu_o))
    assert abs(e_g - e_o) / abs(e_o) < 1e-6, (e_g, e_o)
    D = H.like()
    eng.autodiff_matrix(c, D)
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < 1e-4


def test_energy_scale_matches_oracle(oracle):
    from fastkpm_amd import lattice
    H = lattice.lattice((8, 8), b=2, J=2.0, seed=2)
    n = H.n_rows * H.b_len
    rs = np.random.default_rng(5)
    v = rs.standard_normal(n) + 1j * rs.standard_normal(n)
    eng = _engine()
    es = eng.energy_scale(H, 0.01, 64, v_start=v)
    lo, hi, _, _ = oracle.energy_scale(H, 0.01, 64, v)
    assert abs(es.lo - lo) < 1e-8 * abs(lo) and abs(es.hi - hi) < 1e-8 * abs(hi)
    lam = np.linalg.eigvalsh(H.to_dense())
    assert es.lo <= lam[0] and es.hi >= lam[-1]


@pytest.mark.parametrize("shape,b,ob,mfma,sweep", [((4, 4, 4), 2, False, False, 1),
                                                  ((6, 6, 4), 2, False, False, 8),
                                                  ((5, 3), 1, False, False, 2),
                                                  ((3, 3, 2), 8, True, True, 1),
                                                  ((3, 3, 4), 8, True, True, 8),
                                                  ((3, 3, 2), 8, True, False, 2)])
def test_one_front_sweep_matches_oracle(oracle, shape, b, ob, mfma, sweep):
    """All workgroups sweeping one front over the rows (instead of per-XCD
    regions), in segments of 1, 2 or 8 passes per XCD: vector-unit and
    matrix-core kernels against the oracle."""
    from fastkpm_amd import lattice, EnergyScale
    H = lattice.lattice(shape, b=b, J=1.0 if b == 2 else 0.0, W=1.0 if b < 8 else 0.0,
                        orbital_blocks=ob, seed=13)
    groups = lattice.colouring(shape, tuple(min(2, L) for L in shape), b)
    R = seeded_R(oracle, groups, 17)
    lam = np.linalg.eigvalsh(H.to_dense())
    lo, hi = lam[0] - 0.1, lam[-1] + 0.1
    eng = _engine()
    eng.set_mfma(mfma)
    eng.set_sweep(sweep)
    eng.set_R(R)
    eng.set_H(H, EnergyScale(lo, hi))
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    M = 40
    assert rel_err(eng.moments(M), orc.moments(M)) < FP64_TOL
    for kind in (0, 2):
        assert eng.kernel_config(kind)["sweep"] == sweep
        assert eng.kernel_config(kind)["mfma"] is mfma
    c = fermi_coeffs(oracle, M, lo, hi, kT=0.05)
    D = H.like()
    eng.autodiff_matrix(c, D)
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < FP64_TOL
    eng.stoch_matrix(c, D)
    assert rel_err(D.val, orc.stoch_matrix(c, D)) < FP64_TOL


@pytest.mark.parametrize("store", [True, False])
@pytest.mark.parametrize("name,shape,b,J,W,ob,periods,real", [CASES[2], CASES[3], CASES[4]],
                         ids=[CASES[2][0], CASES[3][0], CASES[4][0]])
def test_autodiff_alpha_table_and_recompute(oracle, store, name, shape, b, J, W, ob, periods,
                                            real):
    """autodiff_matrix with the forward vectors kept (alpha_{m-1} read) and
    with alpha recomputed backward, on H's pattern and on a foreign one."""
    from fastkpm_amd import lattice
    H, R = _case(oracle, shape, b, seed=19, J=J, W=W, orbital_blocks=ob, periods=periods, real=real)
    lam = np.linalg.eigvalsh(H.to_dense())
    lo, hi = lam[0] - 0.1, lam[-1] + 0.1
    M = 48
    eng, orc, D = _run_pair(oracle, H, R, lo, hi, M)
    eng.set_store_alpha(store)
    c = fermi_coeffs(oracle, M, lo, hi, kT=0.05)
    nb = H.n_rows
    Dd = lattice.coo_to_bsr(nb, b, list(range(nb)), list(range(nb)), np.zeros((nb, b, b)))
    for DD in (D, Dd):
        mu = eng.moments(M)
        assert eng.alpha_table() is store
        assert rel_err(mu, orc.moments(M)) < FP64_TOL
        eng.autodiff_matrix(c, DD)
        assert rel_err(DD.val, orc.autodiff_matrix(c, DD)) < FP64_TOL, (name, store)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name,shape,b,J,W,ob,periods,real", [CASES[0], CASES[2], CASES[4]],
                         ids=[CASES[k][0] for k in (0, 2, 4)])
def test_stoch_matrix_from_table(oracle, prec, name, shape, b, J, W, ob, periods, real):
    """stoch_matrix from the forward vectors a preceding moments(M) kept
    (xi = v + 2 T_K(Hs) u), for M' <= M coefficients, odd and tiny M'; M' > M,
    a new H and an explicit Clenshaw request fall back to the recurrence."""
    from fastkpm_amd import EnergyScale
    H, R = _case(oracle, shape, b, seed=23, J=J, W=W, orbital_blocks=ob, periods=periods, real=real)
    lam = np.linalg.eigvalsh(H.to_dense())
    lo, hi = lam[0] - 0.1, lam[-1] + 0.1
    M = 48
    dtype = "double" if real else "cx_double"
    eng, orc, D = _run_pair(oracle, H, R, lo, hi, M, dtype=dtype, prec=prec)
    eng.set_store_alpha(2)
    tol = FP64_TOL if prec == "fp64" else 1e-4
    c = fermi_coeffs(oracle, 2 * M, lo, hi, kT=0.05, kind=1)
    eng.moments(M)
    assert eng.alpha_table()
    for m in (M, M - 1, 30, 3, 2):
        eng.stoch_matrix(c[:m], D)
        assert eng.stoch_method() == "table", m
        assert rel_err(D.val, orc.stoch_matrix(c[:m], D)) < tol, (name, m)
    # autodiff_matrix still needs its moments(M); the table serves both
    eng.moments(M)
    orc.moments(M)
    cE = fermi_coeffs(oracle, M, lo, hi, kT=0.05)
    eng.autodiff_matrix(cE, D)
    assert rel_err(D.val, orc.autodiff_matrix(cE, D)) < tol
    eng.stoch_matrix(c[:M], D)
    assert eng.stoch_method() == "table"
    eng.stoch_matrix(c[:M + 6], D)  # beyond alpha_{M/2}
    assert eng.stoch_method() == "clenshaw"
    assert rel_err(D.val, orc.stoch_matrix(c[:M + 6], D)) < tol
    eng.set_stoch_method("clenshaw")
    eng.stoch_matrix(c[:M], D)
    assert eng.stoch_method() == "clenshaw"
    eng.set_stoch_method("auto")
    eng.set_H(H, EnergyScale(lo - 0.05, hi))  # another Hs: the table is stale
    orc.set_H(H, lo - 0.05, hi)
    eng.stoch_matrix(c[:M], D)
    assert eng.stoch_method() == "clenshaw"
    assert rel_err(D.val, orc.stoch_matrix(c[:M], D)) < tol


def test_option_refusals():
    from fastkpm_amd._capi import FkpmError, OPT_SWEEP, OPT_MFMA, call
    eng = _engine()
    for opt, val in ((OPT_SWEEP, 65), (OPT_SWEEP, -2), (OPT_MFMA, 3), (99, 0)):
        with pytest.raises(FkpmError):
            call("fkpm_set_option", eng.handle, opt, val)


def _ring_no_diagonal(nb=6):
    """A ring of 2x2 hopping blocks with no on-site blocks (the Lanczos
    pattern then gains explicit zero diagonal blocks)."""
    from fastkpm_amd import lattice
    rows, cols, blks = [], [], []
    T = -np.array([[1.0, 0.5j], [-0.5j, 2.0]])
    for i in range(nb):
        j = (i + 1) % nb
        rows += [i, j]
        cols += [j, i]
        blks += [T, T.conj().T]
    return lattice.coo_to_bsr(nb, 2, rows, cols, np.array(blks))


@pytest.mark.parametrize("case", ["invariant_subspace", "no_diagonal", "real", "cx_float"])
def test_energy_scale_cases(oracle, case):
    """Lanczos on the device, coefficients kept there until the end: an
    iteration count beyond n (the recursion meets an invariant subspace and
    stops at the exact extremes), H without diagonal blocks, real and single
    precision engines."""
    from fastkpm_amd import HipEngine, lattice
    if case == "no_diagonal":
        H, dt, iters = _ring_no_diagonal(), "cx_double", 40
    elif case == "real":
        H, dt, iters = lattice.lattice((6, 5), b=1, W=1.0, seed=3, dtype=np.float64), "double", 20
    elif case == "cx_float":
        H, dt, iters = lattice.lattice((4, 4), b=2, J=1.0, seed=3), "cx_float", 24
    else:
        H, dt, iters = lattice.lattice((3, 3), b=2, J=1.5, seed=4), "cx_double", 40
    n = H.n_rows * H.b_len
    rs = np.random.default_rng(7)
    v = rs.standard_normal(n)
    if dt != "double":
        v = v + 1j * rs.standard_normal(n)
    eng = HipEngine(dt, "fp64")
    es = eng.energy_scale(H, 0.0, iters, v_start=v)
    lo, hi, _, _ = oracle.energy_scale(H, 0.0, iters, v)
    tol = 1e-5 if dt == "cx_float" else 1e-9
    assert abs(es.lo - lo) < tol * abs(lo) and abs(es.hi - hi) < tol * abs(hi)
    lam = np.linalg.eigvalsh(H.to_dense())
    if iters >= n:
        assert abs(es.lo - lam[0]) < 1e-9 * abs(lam[0]) and abs(es.hi - lam[-1]) < 1e-9 * abs(lam[-1])


def test_autodiff_requires_moments(oracle):
    from fastkpm_amd import lattice, EnergyScale
    from fastkpm_amd._capi import FkpmError
    H = lattice.lattice((4, 4), b=1)
    eng = _engine()
    eng.set_R(np.eye(16))
    eng.set_H(H, EnergyScale(-4.1, 4.1))
    with pytest.raises(FkpmError):
        eng.autodiff_matrix(np.ones(8), H.like())
    eng.moments(8)
    with pytest.raises(FkpmError):  # different M
        eng.autodiff_matrix(np.ones(10), H.like())


def _pair(oracle, shape=(8, 8), b=2, seed=3, uncorrelated_S=None, **kw):
    from fastkpm_amd import lattice, EnergyScale
    H = lattice.lattice(shape, b=b, seed=seed, **kw)
    n = H.n_rows * b
    if uncorrelated_S:
        R = oracle.set_R_uncorrelated(n, uncorrelated_S, oracle.MT19937(seed))
    else:
        groups = lattice.colouring(shape, tuple(2 for _ in shape), b)
        R = seeded_R(oracle, groups, seed)
    lam = np.linalg.eigvalsh(H.to_dense())
    lo, hi = lam[0] - 0.2, lam[-1] + 0.3
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    return H, R, EnergyScale(lo, hi), orc


@pytest.mark.parametrize("method", ["clenshaw", "forward"])
@pytest.mark.parametrize("dense_R", [False, True])
def test_stoch_matrix_methods(oracle, method, dense_R):
    """Both stoch_matrix formulations, with compact (probing) and dense R."""
    H, R, es, orc = _pair(oracle, J=2.0, W=1.0, uncorrelated_S=12 if dense_R else None)
    eng = _engine()
    eng.set_stoch_method(method)
    eng.set_R(R)
    eng.set_H(H, es)
    M = 90
    c = fermi_coeffs(oracle, M, es.lo, es.hi, kT=0.05, kind=1)
    D = H.like()
    eng.stoch_matrix(c, D)
    assert eng.state()["r_compact"] is (not dense_R)
    assert rel_err(D.val, orc.stoch_matrix(c, D)) < FP64_TOL
    # M = 2 and M = 3 (virtual Clenshaw terms only)
    for m in (2, 3):
        eng.stoch_matrix(c[:m], D)
        assert rel_err(D.val, orc.stoch_matrix(c[:m], D)) < FP64_TOL, m


def test_every_compiled_geometry(oracle):
    """Each compiled (lanes, packets) geometry of b = 2 gives the oracle's results."""
    from fastkpm_amd._capi import FkpmError
    H, R, es, orc = _pair(oracle, J=1.0)
    M = 40
    c = fermi_coeffs(oracle, M, es.lo, es.hi)
    mu_o = orc.moments(M)
    d_o = orc.autodiff_matrix(c, H)
    s_o = orc.stoch_matrix(c, H)
    geos = [(4, 1), (4, 2), (8, 2), (16, 2), (16, 4), (32, 4), (64, 1), (64, 2), (64, 4)]
    for L, KQ in geos:
        for maxwg in (0, 7):
            eng = _engine()
            eng.set_R(R)
            eng.set_H(H, es)
            eng.set_geometry(L, KQ, maxwg)
            assert rel_err(eng.moments(M), mu_o) < FP64_TOL, (L, KQ)
            D = H.like()
            eng.autodiff_matrix(c, D)
            assert rel_err(D.val, d_o) < FP64_TOL, (L, KQ)
            eng.stoch_matrix(c, D)
            assert rel_err(D.val, s_o) < FP64_TOL, (L, KQ)
            eng.close()
    eng = _engine()
    with pytest.raises(FkpmError):
        eng.set_R(R)
        eng.set_H(H, es)
        eng.set_geometry(8, 1, 0)  # not compiled for b = 2


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_column_slices_outer_loop(oracle, prec):
    """FKPM_OPT_SLICE 1: every step kernel sweeps its rows once per column
    pass (S = 40 probe columns: several passes for every geometry below), for
    the forward step, the Clenshaw and forward-order stoch_matrix, and the
    autodiff step with alpha recomputed and read from the kept table."""
    H, R, es, orc = _pair(oracle, J=1.0, uncorrelated_S=40)
    M = 40
    c = fermi_coeffs(oracle, M, es.lo, es.hi)
    mu_o = orc.moments(M)
    d_o = orc.autodiff_matrix(c, H)
    s_o = orc.stoch_matrix(c, H)
    tol_mu, tol_d = (FP64_TOL, FP64_TOL) if prec == "fp64" else (2e-5, 1e-4)
    for L, KQ in [(4, 1), (8, 2), (16, 2), (64, 1)]:
        for store, method in ((0, "clenshaw"), (2, "auto"), (0, "forward")):
            eng = _engine("cx_double", prec)
            eng.set_R(R)
            eng.set_H(H, es)
            eng.set_geometry(L, KQ, 0)
            eng.set_slice(1)
            eng.set_store_alpha(store)
            eng.set_stoch_method(method)
            assert rel_err(eng.moments(M), mu_o) < tol_mu, (L, KQ, store)
            assert eng.kernel_config(0)["nt"] & 4
            D = H.like()
            eng.autodiff_matrix(c, D)
            assert rel_err(D.val, d_o) < tol_d, (L, KQ, store)
            eng.stoch_matrix(c, D)
            assert rel_err(D.val, s_o) < tol_d, (L, KQ, method)
            eng.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lattice_row_order(oracle, prec):
    """FKPM_OPT_ROW_ORDER 1 on a periodic 3-D lattice (16 x 8 x 8 sites, 2x2
    blocks): the step kernels visit the rows slab by slab (one x-slab of two
    sites per XCD) in y-tiles along z.  Moments, dE/dH and stoch_matrix (the
    Clenshaw form, and the forward form with the kept table) against the
    oracle, and against row order: D bitwise the same (a row's arithmetic does
    not depend on when it is visited), the moments up to the order of the
    dot-product sums."""
    H, R, es, orc = _pair(oracle, shape=(16, 8, 8), J=1.0, W=1.0)
    M = 40
    c = fermi_coeffs(oracle, M, es.lo, es.hi)
    mu_o, d_o, s_o = orc.moments(M), orc.autodiff_matrix(c, H), orc.stoch_matrix(c, H)
    tol, tol_d = (FP64_TOL, FP64_TOL) if prec == "fp64" else (2e-5, 1e-4)
    for L, KQ in [(64, 1), (64, 2), (16, 2), (4, 1)]:
        out = {}
        for order in (0, 1):
            for store, method in ((0, "clenshaw"), (2, "forward")):
                eng = _engine("cx_double", prec)
                eng.set_R(R)
                eng.set_H(H, es)
                eng.set_geometry(L, KQ, 0)
                eng.set_row_order(order)
                eng.set_store_alpha(store)
                eng.set_stoch_method(method)
                mu = eng.moments(M)
                D = H.like()
                eng.autodiff_matrix(c, D)
                S = H.like()
                eng.stoch_matrix(c, S)
                for k in (0, 1, 2, 4, 5):
                    kc = eng.kernel_config(k)
                    assert bool(kc["nt"] & 8) is bool(order), (k, kc)
                    if order:
                        assert kc["sweep"] == 0
                assert not eng.kernel_config(3)["nt"] & 8
                eng.close()
                assert rel_err(mu, mu_o) < tol, (L, KQ, order, store)
                assert rel_err(D.val, d_o) < tol_d, (L, KQ, order, store)
                assert rel_err(S.val, s_o) < tol_d, (L, KQ, order, method)
                out[order, store] = (mu, D.val.copy(), S.val.copy())
        for store in (0, 2):
            assert np.array_equal(out[0, store][1], out[1, store][1]), (L, KQ, store)
            assert np.array_equal(out[0, store][2], out[1, store][2]), (L, KQ, store)
            assert rel_err(out[0, store][0], out[1, store][0]) < (1e-13 if prec == "fp64" else 1e-6)


def test_row_order_needs_a_3d_lattice(oracle):
    """A 2-D lattice has no row order: the switch leaves row order in place."""
    H, R, es, orc = _pair(oracle, shape=(16, 16), J=1.0)
    eng = _engine()
    eng.set_R(R)
    eng.set_H(H, es)
    eng.set_row_order(1)
    mu = eng.moments(20)
    assert not eng.kernel_config(0)["nt"] & 8
    assert rel_err(mu, orc.moments(20)) < FP64_TOL
    eng.close()


@pytest.mark.parametrize("prec,store", [("fp64", 1), ("fp32", 1), ("fp64", 2), ("fp32", 2)])
def test_autotuned_engine(oracle, prec, store):
    """store 2: the forward vectors are kept, and the tuner also times the
    table-writing forward step and the table-reading backward step."""
    H, R, es, orc = _pair(oracle, shape=(16, 16), J=2.0)
    eng = _engine("cx_double", prec)
    eng.set_store_alpha(store)
    eng.set_R(R)
    eng.set_H(H, es)
    eng.autotune()
    assert eng.state()["tuned"]
    M = 120
    c = fermi_coeffs(oracle, M, es.lo, es.hi)
    mu = eng.moments(M)
    assert eng.alpha_table() is (store == 2)
    D = H.like()
    eng.autodiff_matrix(c, D)
    S = H.like()
    eng.stoch_matrix(c, S)
    tol = FP64_TOL if prec == "fp64" else 2e-5
    assert rel_err(mu, orc.moments(M)) < tol
    assert rel_err(D.val, orc.autodiff_matrix(c, D)) < (tol if prec == "fp64" else 1e-4)
    assert rel_err(S.val, orc.stoch_matrix(c, S)) < (tol if prec == "fp64" else 1e-4)
    # a second engine of the same shape reuses the tuned geometry
    eng2 = _engine("cx_double", prec)
    eng2.set_store_alpha(store)
    eng2.set_R(R)
    eng2.set_H(H, es)
    eng2.moments(4)
    assert eng2.state()["tuned"]
    for k in range(6):
        assert eng2.kernel_config(k) == eng.kernel_config(k)


def test_bandwidth_probe():
    """The copy-bandwidth probe: a finite rate for a ragged size, refusals."""
    from fastkpm_amd import bandwidth_probe
    from fastkpm_amd._capi import FkpmError
    eng = _engine()
    g = bandwidth_probe(eng, (64 << 20) + 16 * 1000, 3)  # not a multiple of a workgroup's span
    assert np.isfinite(g) and g > 50.0, g
    for nbytes, reps in ((1000 * 16 + 8, 1), (0, 1), (1 << 20, 0)):
        with pytest.raises(FkpmError):
            bandwidth_probe(eng, nbytes, reps)


@pytest.mark.parametrize("env", [None, "1"])
def test_autotune_from_environment(oracle, monkeypatch, env):
    """FKPM_AUTOTUNE=1 at open tunes a shape at its first use (the switch for
    callers without an option call); without it the default configuration
    runs.  Shapes of their own, so that no earlier test tuned them."""
    if env is None:
        monkeypatch.delenv("FKPM_AUTOTUNE", raising=False)
        shape = (10, 6)
    else:
        monkeypatch.setenv("FKPM_AUTOTUNE", env)
        shape = (12, 6)
    H, R, es, orc = _pair(oracle, shape=shape, J=1.0)
    eng = _engine()
    eng.set_R(R)
    eng.set_H(H, es)
    M = 40
    mu = eng.moments(M)
    assert eng.state()["tuned"] is (env is not None)
    assert rel_err(mu, orc.moments(M)) < FP64_TOL
    assert eng.kernel_config(0)["sweep"] in ((2,) if env is None else (0, 1, 2, 8))
