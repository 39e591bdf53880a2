"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

No reference fixtures exist for this path (the reference has no tests), so each
oracle function is pinned by what fixes it independently: the libstdc++
implementation of the random engine and distribution (committed fixture), the
C++ standard's known answer, closed forms of Chebyshev / Jackson / Fermi
functions, exact diagonalisation (numpy) with R = identity, and central finite
differences of the energy for the reverse-mode gradient.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def _lat(shape, b=1, **kw):
    from fastkpm_amd import lattice
    return lattice.lattice(shape, b=b, **kw)


# -- random numbers -----------------------------------------------------------

def test_mt19937_known_answer(oracle):
    g = oracle.MT19937()  # default seed 5489
    for _ in range(9999):
        g()
    assert g() == 4123659995  # C++ standard [rand.predef]


def test_rng_matches_libstdcxx_fixture(oracle):
    fx = json.loads((GOLDEN / "rng_libstdcxx.json").read_text())
    g = oracle.MT19937()
    assert [g() for _ in range(8)] == fx["mt19937_default_first"]
    for case in fx["uniform_int"]:
        g = oracle.MT19937(case["seed"])
        draws = [g.uniform_int(case["a"], case["b"]) for _ in range(len(case["draws"]))]
        assert draws == case["draws"], case
        assert g() == case["next_raw"], case  # same number of raw draws consumed


def test_R_correlated_structure(oracle):
    from fastkpm_amd import lattice
    groups = lattice.colouring((8, 8), (4, 4), 2)
    n, s = len(groups), int(groups.max()) + 1
    g = oracle.MT19937(7)
    R = oracle.set_R_correlated(groups, g)
    assert R.shape == (n, s)
    nz = np.abs(R) > 0
    assert np.all(nz.sum(axis=1) == 1)
    assert np.all(np.argmax(nz, axis=1) == groups)
    assert np.allclose(np.abs(R[nz]), 1.0)
    assert set(np.round(R[nz], 12)) <= {1, -1, 1j, -1j}
    gram = R.conj().T @ R
    assert np.allclose(np.diag(gram), np.bincount(groups, minlength=s))
    # the caller's engine advanced by exactly one draw per column
    h = oracle.MT19937(7)
    h.discard(s)
    assert g() == h()
    # colour draws of all four phases occur (Z4, not Z2)
    assert len(set(np.round(R[nz], 12))) == 4


def test_R_uncorrelated_and_identity(oracle):
    n, s = 50, 9
    g = oracle.MT19937(3)
    R = oracle.set_R_uncorrelated(n, s, g, 2, 7)
    assert R.shape == (n, 5)
    assert np.allclose(np.abs(R), 1 / np.sqrt(s))
    Rr = oracle.set_R_uncorrelated(n, s, oracle.MT19937(3), is_complex=False)
    assert np.all(np.isin(np.round(Rr.real * np.sqrt(s), 12), [-1.0, 1.0]))
    # columns j0..j1 of the full matrix equal the sliced call (column seeds are
    # the j-th master draws)
    full = oracle.set_R_uncorrelated(n, s, oracle.MT19937(3))
    g2 = oracle.MT19937(3)
    g2.discard(2)
    part = oracle.set_R_uncorrelated(n, s, g2, 2, 7)
    assert np.array_equal(full[:, 2:7], part)
    I = oracle.set_R_identity(6, 1, 4)
    assert np.array_equal(I, np.eye(6)[:, 1:4])
    # two-argument correlated form with s >= n is the identity
    assert np.array_equal(oracle.set_R_correlated(np.arange(5), oracle.MT19937(1)), np.eye(5))


# -- BSR ------------------------------------------------------------------------

def test_bsr_build_sorts_and_sums_duplicates(oracle):
    rs = np.random.default_rng(0)
    nb, b, nin = 7, 3, 40
    rows = rs.integers(0, nb, nin)
    cols = rs.integers(0, nb, nin)
    blocks = rs.standard_normal((nin, b, b)) + 1j * rs.standard_normal((nin, b, b))
    vals = np.transpose(blocks, (0, 2, 1)).reshape(-1)  # column-major blocks
    rp, ri, ci, val = oracle.bsr_build(nb, b, rows, cols, vals)
    dense = np.zeros((nb * b, nb * b), complex)
    for r, c, blk in zip(rows, cols, blocks):
        dense[b * r:b * r + b, b * c:b * c + b] += blk
    got = np.zeros_like(dense)
    for k in range(len(ri)):
        got[b * ri[k]:b * ri[k] + b, b * ci[k]:b * ci[k] + b] += val[k * b * b:(k + 1) * b * b].reshape(b, b).T
    assert np.allclose(got, dense, atol=1e-14)
    keys = list(zip(ri, ci))
    assert keys == sorted(set(keys))
    assert list(rp) == [int(np.sum(ri < i)) for i in range(nb + 1)]


def test_symmetrize_is_hermitian_mean(oracle):
    from fastkpm_amd import lattice
    H = lattice.lattice((4, 4), b=2, J=1.0, seed=2)
    rs = np.random.default_rng(1)
    v = rs.standard_normal(len(H.val)) + 1j * rs.standard_normal(len(H.val))
    D = H.like()
    D.val = v
    A = D.to_dense()
    out = oracle.bsr_symmetrize(2, D.row_ptr, D.row_idx, D.col_idx, v)
    D.val = out
    S = D.to_dense()
    mask = np.abs(H.to_dense()) + np.eye(A.shape[0]) > 0
    assert np.allclose(S[mask], ((A + A.conj().T) / 2)[mask], atol=1e-15)
    # a lower block (i > j) without its partner is rejected (the reference
    # aborts in find_index); a lone upper block (i < j) is left unchanged
    with pytest.raises(ValueError):
        oracle.bsr_symmetrize(1, np.array([0, 1, 3]), np.array([0, 1, 1]), np.array([0, 0, 1]),
                              np.ones(3, complex))
    out = oracle.bsr_symmetrize(1, np.array([0, 2, 2]), np.array([0, 0]), np.array([0, 1]),
                                np.array([1 + 2j, 3 - 1j]))
    assert np.array_equal(out, np.array([1 + 0j, 3 - 1j]))


# -- spectral closed forms ---------------------------------------------------------

def test_jackson_kernel_closed_forms(oracle):
    for M in (4, 37, 500):
        g = oracle.jackson_kernel(M)
        assert g[0] == pytest.approx(1.0, abs=1e-15)
        assert g[1] == pytest.approx(np.cos(np.pi / (M + 1)), rel=1e-14)
        assert np.all(np.diff(g) < 0) and np.all(g > 0)


def test_chebyshev_fill_array(oracle):
    t = 0.731
    T = oracle.chebyshev_fill_array(np.cos(t), 40, 1)
    U = oracle.chebyshev_fill_array(np.cos(t), 40, 2)
    m = np.arange(40)
    assert np.allclose(T, np.cos(m * t), atol=1e-13)
    assert np.allclose(U, np.sin((m + 1) * t) / np.sin(t), atol=1e-12)


def test_expansion_coefficients_exact_polynomials(oracle):
    M, Mq, lo, hi = 24, 48, -3.0, 5.0
    # f = 1: c = (1, 0, 0, ...)
    c = oracle.expansion_coefficients_fermi(M, Mq, oracle.CONSTANT_ONE, 0, 0, lo, hi)
    assert c[0] == pytest.approx(1.0, abs=1e-14) and np.max(np.abs(c[1:])) < 1e-14
    # f = T_0 + 0.5 T_2 - 0.25 T_7 of the scaled variable: c_m = g_m * coefficient
    x = np.cos(np.pi * (np.arange(Mq) + 0.5) / Mq)
    f = 1 + 0.5 * np.cos(2 * np.arccos(x)) - 0.25 * np.cos(7 * np.arccos(x))
    c = oracle.expansion_coefficients_samples(M, Mq, f)
    g = oracle.jackson_kernel(M)
    want = np.zeros(M)
    want[0], want[2], want[7] = 1.0, 0.5 * g[2], -0.25 * g[7]
    assert np.allclose(c, want, atol=1e-14)
    # the sample points are the Chebyshev nodes mapped to [lo, hi]
    xs = oracle.chebyshev_points(Mq, lo, hi)
    assert np.allclose(xs, (x * (hi - lo) / 2 + (hi + lo) / 2))


def test_fermi_functions(oracle):
    assert oracle.fermi_energy(-1.5, 0.0, 0.2) == pytest.approx(-1.7)
    assert oracle.fermi_energy(0.5, 0.0, 0.2) == 0.0
    assert oracle.fermi_density(-1.0, 0.0, 0.0) == 1.0 and oracle.fermi_density(1.0, 0.0, 0.0) == 0.0
    kT, mu = 0.05, 0.1
    assert oracle.fermi_energy(mu, kT, mu) == pytest.approx(-kT * np.log(2), rel=1e-14)
    assert oracle.fermi_density(mu, kT, mu) == pytest.approx(0.5)
    for x in (-0.2, 0.05, 0.13, 0.4):  # d/dx fermi_energy = fermi_density
        h = 1e-6
        d = (oracle.fermi_energy(x + h, kT, mu) - oracle.fermi_energy(x - h, kT, mu)) / (2 * h)
        assert d == pytest.approx(oracle.fermi_density(x, kT, mu), abs=1e-7)


def test_moment_product_equals_density_product(oracle):
    """sum c_m mu_m == (1/Mq) sum gamma_i f(x_i) with gamma the transform of mu:
    two independent code paths (fastkpm.cpp:41-78 vs 236-256, 314-322)."""
    rs = np.random.default_rng(4)
    M, Mq, lo, hi, kT, mu0 = 64, 128, -4.5, 3.9, 0.07, -0.3
    mu = rs.standard_normal(M) / (1 + np.arange(M))
    c = oracle.expansion_coefficients_fermi(M, Mq, oracle.FERMI_ENERGY, kT, mu0, lo, hi)
    e1 = oracle.moment_product(c, mu)
    e2 = oracle.electronic_grand_energy(oracle.moment_transform(mu, Mq), lo, hi, kT, mu0)
    assert e1 == pytest.approx(e2, rel=1e-12)


# -- the engine ---------------------------------------------------------------------

def _dense_moments(Hd, lo, hi, M):
    lam = np.linalg.eigvalsh(Hd)
    x = (lam - (hi + lo) / 2) / ((hi - lo) / 2)
    return np.array([np.cos(m * np.arccos(x)).sum() for m in range(M)])


def test_moments_identity_R_exact(oracle):
    for H, lo, hi in ((_lat((6, 6), 1, W=1.5, seed=1), -4.9, 5.3),
                      (_lat((4, 4), 2, J=2.0, seed=3), -6.5, 6.1),
                      (_lat((3, 3), 4, orbital_blocks=True, seed=5), -9.0, 9.5)):
        n = H.n_rows * H.b_len
        e = oracle.Engine()
        e.set_H(H, lo, hi)
        e.set_R(np.eye(n))
        M = 30
        assert np.allclose(e.moments(M), _dense_moments(H.to_dense(), lo, hi, M), atol=1e-11 * n)


def test_moments_periodic_square_closed_form(oracle):
    L, t, lo, hi = 8, -1.0, -4.1, 4.1
    H = _lat((L, L), 1, t=t)
    e = oracle.Engine()
    e.set_H(H, lo, hi)
    e.set_R(np.eye(L * L))
    k = 2 * np.pi * np.arange(L) / L
    eps = (2 * t * (np.cos(k)[:, None] + np.cos(k)[None, :])).reshape(-1)
    x = eps / 4.1
    want = np.array([np.cos(m * np.arccos(x)).sum() for m in range(40)])
    assert np.allclose(e.moments(40), want, atol=1e-11)


def test_moment_column_shards_sum(oracle):
    """moments over S-column shards add up (the Engine_Threaded split)."""
    from fastkpm_amd import lattice
    H = _lat((6, 6), 2, J=1.0, seed=8)
    groups = lattice.colouring((6, 6), (3, 3), 2)
    R = oracle.set_R_correlated(groups, oracle.MT19937(5))
    e = oracle.Engine()
    e.set_H(H, -6.0, 6.0)
    e.set_R(R)
    full = e.moments(20)
    a = e.moments_cols(20, 0, 7)
    b = e.moments_cols(20, 7, R.shape[1])
    assert np.allclose(a + b, full, atol=1e-12)


def test_stoch_matrix_identity_R_is_p_of_H(oracle):
    H = _lat((4, 4), 2, J=1.7, seed=6)
    lo, hi = -6.3, 6.0
    n = H.n_rows * 2
    M = 40
    c = oracle.expansion_coefficients_fermi(M, 2 * M, oracle.FERMI_DENSITY, 0.1, 0.2, lo, hi)
    e = oracle.Engine()
    e.set_H(H, lo, hi)
    e.set_R(np.eye(n))
    D = H.like()
    D.val = e.stoch_matrix(c, D)
    lam, V = np.linalg.eigh(H.to_dense())
    x = (lam - (hi + lo) / 2) / ((hi - lo) / 2)
    p = np.array([sum(c[m] * np.cos(m * np.arccos(xx)) for m in range(M)) for xx in x])
    P = (V * p) @ V.conj().T
    mask = np.abs(H.to_dense()) > 0
    Dd = D.to_dense()
    assert np.allclose(Dd[mask], P[mask], atol=1e-12)


def _energy(oracle, H, R, lo, hi, c):
    e = oracle.Engine()
    e.set_H(H, lo, hi)
    e.set_R(R)
    return float(np.dot(c, e.moments(len(c))))


def test_autodiff_matches_finite_differences(oracle):
    """D from autodiff_matrix is dE/dH^T of E(H) = sum c_m mu_m(H) at fixed R and
    energy scale: (E(H+eps Delta) - E(H-eps Delta)) / 2 eps = Re tr(D Delta)."""
    from fastkpm_amd import lattice
    H = _lat((4, 4), 2, J=2.0, seed=9)
    lo, hi = -6.6, 6.4
    groups = lattice.colouring((4, 4), (2, 2), 2)
    R = oracle.set_R_correlated(groups, oracle.MT19937(11))
    M = 24
    c = oracle.expansion_coefficients_fermi(M, 2 * M, oracle.FERMI_ENERGY, 0.2, 0.1, lo, hi)
    e = oracle.Engine()
    e.set_H(H, lo, hi)
    e.set_R(R)
    e.moments(M)
    D = H.like()
    D.val = e.autodiff_matrix(c, D)
    Dd = D.to_dense()
    rs = np.random.default_rng(2)
    b = 2
    for (i, j) in ((0, 0), (0, 1), (5, 9), (3, 7)):
        k = H.find_index(i, j)
        delta = rs.standard_normal((b, b)) + 1j * rs.standard_normal((b, b))
        if i == j:
            delta = (delta + delta.conj().T) / 2
        Dl = np.zeros((H.n_rows * b,) * 2, complex)
        Dl[b * i:b * i + b, b * j:b * j + b] += delta
        if i != j:
            Dl[b * j:b * j + b, b * i:b * i + b] += delta.conj().T
        eps = 1e-4
        Hp, Hm = H.like(), H.like()
        Hp.val = H.val.copy()
        Hm.val = H.val.copy()
        for (r, cc, blk) in ((i, j, delta),) + (((j, i, delta.conj().T),) if i != j else ()):
            kk = H.find_index(r, cc)
            Hp.val[kk * 4:(kk + 1) * 4] += eps * blk.T.reshape(-1)
            Hm.val[kk * 4:(kk + 1) * 4] -= eps * blk.T.reshape(-1)
        fd = (_energy(oracle, Hp, R, lo, hi, c) - _energy(oracle, Hm, R, lo, hi, c)) / (2 * eps)
        an = np.real(np.trace(Dd @ Dl))
        assert an == pytest.approx(fd, rel=1e-6, abs=1e-9), (i, j, k)


def test_autodiff_identity_R_is_derivative(oracle):
    """With R = I, D = p'(Hs)/mag with p' = sum c_m m U_{m-1} (on H's pattern)."""
    H = _lat((3, 3), 2, J=1.2, seed=4)
    lo, hi = -6.0, 5.5
    n = 18
    M = 20
    c = oracle.expansion_coefficients_fermi(M, 2 * M, oracle.FERMI_ENERGY, 0.3, -0.2, lo, hi)
    e = oracle.Engine()
    e.set_H(H, lo, hi)
    e.set_R(np.eye(n))
    e.moments(M)
    D = H.like()
    D.val = e.autodiff_matrix(c, D)
    lam, V = np.linalg.eigh(H.to_dense())
    mag = (hi - lo) / 2
    x = (lam - (hi + lo) / 2) / mag
    th = np.arccos(x)
    dp = np.array([sum(c[m] * m * np.sin(m * t) / np.sin(t) for m in range(1, M)) for t in th])
    P = (V * dp) @ V.conj().T / mag
    mask = np.abs(H.to_dense()) > 0
    assert np.allclose(D.to_dense()[mask], P[mask], atol=1e-11)


def test_autodiff_requires_moments_with_same_M(oracle):
    H = _lat((3, 3), 1)
    e = oracle.Engine()
    e.set_H(H, -4.1, 4.1)
    e.set_R(np.eye(9))
    with pytest.raises(ValueError):
        e.autodiff_matrix(np.ones(6), H.like())
    e.moments(8)
    with pytest.raises(ValueError):
        e.autodiff_matrix(np.ones(6), H.like())


def test_tridiagonal_extremes_and_lanczos(oracle):
    rs = np.random.default_rng(7)
    a = rs.standard_normal(40)
    b = np.concatenate([[0.0], np.abs(rs.standard_normal(39))])
    T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    lam = np.linalg.eigvalsh(T)
    lo, hi = oracle.tridiag_extremes(a, b)
    assert lo == pytest.approx(lam[0], abs=1e-12) and hi == pytest.approx(lam[-1], abs=1e-12)
    # Lanczos with as many iterations as the dimension finds the exact extremes
    H = _lat((4, 4), 2, J=1.5, seed=2)
    n = 32
    v = rs.standard_normal(n) + 1j * rs.standard_normal(n)
    lo, hi, _, _ = oracle.energy_scale(H, 0.0, n, v)
    lam = np.linalg.eigvalsh(H.to_dense())
    assert lo == pytest.approx(lam[0], abs=1e-8) and hi == pytest.approx(lam[-1], abs=1e-8)
    lo2, hi2, _, _ = oracle.energy_scale(H, 0.1, n, v)
    assert lo2 == pytest.approx(lo - 0.1 * (hi - lo)) and hi2 == pytest.approx(hi + 0.1 * (hi - lo))


def test_moments2_identity_currents_closed_form(oracle):
    """j1 = j2 = I and R = I: mu2[m1][m2] = tr T_m1 T_m2 = (mu_{m1+m2} + mu_{|m1-m2|}) / 2."""
    from fastkpm_amd import lattice
    H = lattice.lattice((3, 3), b=2, J=1.3, seed=2)
    n = H.n_rows * 2
    e = oracle.Engine()
    e.set_H(H, -6.0, 6.0)
    e.set_R(np.eye(n, dtype=complex))
    M = 10
    I = lattice.coo_to_bsr(H.n_rows, 2, list(range(H.n_rows)), list(range(H.n_rows)),
                           np.array([np.eye(2)] * H.n_rows))
    mu2 = e.moments2_v1(M, I, I)
    mu = e.moments(2 * M)
    ref = np.array([[0.5 * (mu[a + b] + mu[abs(a - b)]) for b in range(M)] for a in range(M)])
    assert np.max(np.abs(mu2 - ref)) < 1e-12 * np.max(np.abs(ref))


def test_moments2_matches_dense_traces(oracle):
    """mu2[m1][m2] = tr(R^+ T_m1(Hs)^+ j2 T_m2(Hs) j1 R) by dense matrix algebra."""
    from fastkpm_amd import lattice
    H = lattice.lattice((3, 3), b=2, J=1.3, seed=2)
    n = H.n_rows * 2
    lo, hi = -6.0, 6.0
    Hd = H.to_dense()
    Hs = (Hd - (hi + lo) / 2 * np.eye(n)) / ((hi - lo) / 2)
    rs = np.random.default_rng(1)
    J1, J2 = H.like(), H.like()
    J1.val = rs.normal(size=H.val.shape) + 1j * rs.normal(size=H.val.shape)
    J2.val = rs.normal(size=H.val.shape) + 0j
    M = 10
    T = [np.eye(n), Hs]
    for _ in range(2, M):
        T.append(2 * Hs @ T[-1] - T[-2])
    R = rs.normal(size=(n, 3)) + 1j * rs.normal(size=(n, 3))
    e = oracle.Engine()
    e.set_H(H, lo, hi)
    e.set_R(R)
    mu2 = e.moments2_v1(M, J1, J2)
    ref = np.array([[np.trace(R.conj().T @ T[a].conj().T @ J2.to_dense() @ T[b] @ J1.to_dense() @ R)
                     for b in range(M)] for a in range(M)])
    assert np.max(np.abs(mu2 - ref)) < 1e-12 * np.max(np.abs(ref))


def test_lattice_current_is_commutator():
    """lattice.current(H, shape, axis) is Hermitian and, on a lattice too
    long to wrap within one hop's minimum image, equals i [H, X] with X the
    position along the axis unwrapped (bonds across the boundary aside)."""
    from fastkpm_amd import lattice
    shape = (7, 5)
    H = lattice.lattice(shape, b=2, J=1.1, W=0.5, seed=4)
    J = lattice.current(H, shape, 0)
    Jd, Hd = J.to_dense(), H.to_dense()
    assert np.allclose(Jd, Jd.conj().T, atol=1e-14)
    x = np.repeat(np.arange(np.prod(shape)) % shape[0], 2).astype(float)
    comm = 1j * (Hd * x[None, :] - x[:, None] * Hd)
    cross = np.abs(x[:, None] - x[None, :]) > 1  # bonds across the boundary
    assert np.allclose(np.where(cross, 0, Jd), np.where(cross, 0, comm), atol=1e-14)
    assert np.all(np.abs(np.diag(Jd)) == 0)
