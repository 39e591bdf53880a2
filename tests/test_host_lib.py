"""The C++ host half of the product (libfastkpm.so) against the oracle, on the
CPU: probing vectors bit-exact, spectral functions to rounding."""
from __future__ import annotations

import numpy as np
import pytest

DTYPES = [("cx_double", True, False), ("cx_float", True, True), ("double", False, False),
          ("float", False, True)]


@pytest.fixture(scope="module")
def host():
    from fastkpm_amd import host
    return host


def test_rng_is_std_mt19937(host, oracle):
    a, b = host.RNG(123), oracle.MT19937(123)
    assert [a() for _ in range(1000)] == [b() for _ in range(1000)]
    a.discard(777)
    b.discard(777)
    assert a() == b()


@pytest.mark.parametrize("dtype,cx,single", DTYPES)
def test_set_R_correlated_bit_exact(host, oracle, dtype, cx, single):
    from fastkpm_amd import lattice
    groups = lattice.colouring((8, 16), (4, 8), 2)
    r_host = host.RNG(99)
    got = host.set_R_correlated(dtype, groups, r_host)
    r_orc = oracle.MT19937(99)
    want = oracle.set_R_correlated(groups, r_orc, is_complex=cx)
    assert got.shape == want.shape
    assert np.array_equal(got.astype(np.complex128), want)
    assert r_host() == r_orc()  # both engines advanced by one draw per column
    # four-argument form, a column slice
    got = host.set_R_correlated(dtype, groups, host.RNG(5), 3, 11)
    g = oracle.MT19937(5)
    want = oracle.set_R_correlated(groups, g, 3, 11, is_complex=cx)
    assert np.array_equal(got.astype(np.complex128), want)


@pytest.mark.parametrize("dtype,cx,single", DTYPES)
def test_set_R_uncorrelated_and_identity_bit_exact(host, oracle, dtype, cx, single):
    got = host.set_R_uncorrelated(dtype, 77, 13, host.RNG(8), 1, 9)
    want = oracle.set_R_uncorrelated(77, 13, oracle.MT19937(8), 1, 9, is_complex=cx, single=single)
    assert np.array_equal(got.astype(np.complex128), want)
    assert np.array_equal(host.set_R_identity(dtype, 9, 2, 5).astype(np.complex128),
                          oracle.set_R_identity(9, 2, 5))
    # groups with as many colours as rows: the identity branch (engine.cpp:110-115)
    got = host.set_R_correlated(dtype, np.arange(6), host.RNG(1))
    assert np.array_equal(got.astype(np.complex128), np.eye(6))


def test_spectral_functions_match_oracle(host, oracle):
    M, Mq, lo, hi = 300, 600, -6.7, 6.1
    for kind in (host.FERMI_ENERGY, host.FERMI_DENSITY, host.CONSTANT_ONE):
        for kT in (0.0, 0.01, 0.3):
            a = host.expansion_coefficients(M, Mq, kind, kT, 0.25, lo, hi)
            b = oracle.expansion_coefficients_fermi(M, Mq, kind, kT, 0.25, lo, hi)
            assert np.allclose(a, b, rtol=0, atol=1e-15 * max(1.0, np.abs(b).max()))
    assert np.array_equal(host.jackson_kernel(M), oracle.jackson_kernel(M))
    mu = np.random.default_rng(0).standard_normal(M) / (1 + np.arange(M))
    c = oracle.expansion_coefficients_fermi(M, Mq, 0, 0.05, 0.0, lo, hi)
    assert host.moment_product(c, mu) == pytest.approx(oracle.moment_product(c, mu), rel=1e-15)
    g1, g2 = host.moment_transform(mu, Mq), oracle.moment_transform(mu, Mq)
    assert np.allclose(g1, g2, rtol=1e-13, atol=1e-13)
    assert host.electronic_grand_energy(g1, lo, hi, 0.05, 0.1) == pytest.approx(
        oracle.electronic_grand_energy(g2, lo, hi, 0.05, 0.1), rel=1e-12)


def test_filling_to_mu_inverts_filling(host, oracle):
    """filling_to_mu finds mu with mu_to_filling(mu) = filling (bracket width
    2^-30 relative), checked through the oracle's density product."""
    from fastkpm_amd import lattice
    H = lattice.lattice((8, 8), b=1, W=2.0, seed=3)
    lo, hi, M, Mq = -5.2, 5.2, 200, 400
    e = oracle.Engine()
    e.set_H(H, lo, hi)
    e.set_R(np.eye(64))
    mu = e.moments(M)
    gamma = oracle.moment_transform(mu, Mq)
    kT = 0.05
    xs = oracle.chebyshev_points(Mq, lo, hi)
    tot = gamma.sum()
    for filling in (0.2, 0.5, 0.77):
        m = host.filling_to_mu(gamma, lo, hi, kT, filling)
        occ = sum(g * oracle.fermi_density(x, kT, m) for g, x in zip(gamma, xs)) / tot
        assert occ == pytest.approx(filling, abs=1e-6)


def _dct_form_static(oracle, M, Mq, kT, mu, lo, hi, kernel):
    """Static conductivity coefficients from DCT-II / DST-II transforms of two
    node functions (the FFTW branch of src/fastkpm.cpp:90-140), on scipy's
    transforms."""
    from scipy.fft import dct, dst
    avg, mag = (hi + lo) / 2, (hi - lo) / 2
    x = np.cos(np.pi * (np.arange(Mq) + 0.5) / Mq)
    t = 1.0 - x * x
    f = np.array([oracle.fermi_density(xi * mag + avg, kT, mu) for xi in x])
    keep = t >= 1e-4
    x1 = np.where(keep, f * x / np.where(keep, t, 1.0) ** 1.5, 0.0)
    x2 = np.where(keep, f / np.where(keep, t, 1.0), 0.0)
    yc1, yc2, ys1, ys2 = dct(x1, 2), dct(x2, 2), dst(x1, 2), dst(x2, 2)

    def c_(y, k):
        return y[abs(k)]

    def s_(y, k):
        return 0.0 if k == 0 else np.sign(k) * y[abs(k) - 1]

    out = np.zeros((M, M), dtype=np.complex128)
    w = np.where(np.arange(M) == 0, 1.0, 2.0) * kernel
    for a in range(M):
        for b in range(M):
            ms, md = a + b, a - b
            yr = 2 * yc1[ms] + 2 * c_(yc1, md) + (ms * ys2[ms - 1] if ms > 0 else 0.0) \
                + md * s_(ys2, md)
            yi = 2 * s_(ys1, md) - md * (yc2[ms] + c_(yc2, md))
            out[a, b] = np.pi * w[a] * w[b] / (2 * Mq * mag * mag) * (yr + 1j * yi)
    return out


def _dct_form_optical(oracle, M, Mq, kT, mu, omega, lo, hi, kernel):
    """Optical conductivity coefficients, one DCT-II per m1 (the FFTW branch of
    src/fastkpm.cpp:176-205)."""
    from scipy.fft import dct
    avg, mag = (hi + lo) / 2, (hi - lo) / 2
    ws = omega / mag
    i0 = int(np.ceil(np.arccos(1.0 - ws) / np.pi * Mq - 0.5))
    x = np.cos(np.pi * (np.arange(Mq) + 0.5) / Mq)
    out = np.zeros((M, M), dtype=np.complex128)
    for a in range(M):
        X = np.zeros(Mq)
        for i in range(i0, Mq):
            xw = x[i] + ws
            if x[i] < -1 + 1e-4 or xw > 1 - 1e-4:
                continue
            e = x[i] * mag + avg
            df = oracle.fermi_density(e, kT, mu) - oracle.fermi_density(e + omega, kT, mu)
            X[i] = 0.5 * np.cos(a * np.arccos(xw)) * df / (ws * np.sqrt(1 - xw * xw))
        out[a, :] = dct(X, 2)[:M]
    w = np.where(np.arange(M) == 0, 1.0, 2.0) * kernel
    return out * (2 * np.pi * np.outer(w, w) / (Mq * mag * mag))


@pytest.mark.parametrize("kT,mu", [(0.05, 0.0), (0.2, -0.7), (0.0, 0.3)])
def test_conductivity_static_matches_dct_form(host, oracle, kT, mu):
    M, Mq, lo, hi = 24, 96, -3.2, 2.6
    g = host.jackson_kernel(M)
    got = host.conductivity_coefficients(M, Mq, kT, mu, 0.0, lo, hi)
    want = _dct_form_static(oracle, M, Mq, kT, mu, lo, hi, g)
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()
    assert np.abs(want).max() > 0


@pytest.mark.parametrize("omega", [0.3, 1.1])
def test_conductivity_optical_matches_dct_form(host, oracle, omega):
    M, Mq, lo, hi, kT, mu = 20, 200, -3.0, 3.0, 0.1, 0.2
    g = host.jackson_kernel(M)
    got = host.conductivity_coefficients(M, Mq, kT, mu, omega, lo, hi)
    want = _dct_form_optical(oracle, M, Mq, kT, mu, omega, lo, hi, g)
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()
    assert np.abs(got.imag).max() == 0.0


def test_conductivity_edges(host):
    M = 8
    # omega at or beyond the scaled bandwidth: all zero
    assert not np.any(host.conductivity_coefficients(M, 32, 0.1, 0.0, 4.0, -1.0, 1.0))
    with pytest.raises(ValueError, match="Mq"):
        host.conductivity_coefficients(M, 2 * M - 1, 0.1, 0.0, 0.0, -1.0, 1.0)
    with pytest.raises(ValueError, match="omega"):
        host.conductivity_coefficients(M, 32, 0.1, 0.0, -0.1, -1.0, 1.0)
    with pytest.raises(ValueError, match="20 nodes"):
        host.conductivity_coefficients(M, 32, 0.1, 0.0, 1.9, -1.0, 1.0)


def _direct_static(oracle, M, Mq, kT, mu, lo, hi, kernel):
    """Static coefficients as the node sum of the branch without FFTW
    (src/fastkpm.cpp:144-174): sum over nodes of T_m1 (T_m2 - i sin m2 t)
    (x + i m2 sqrt(1 - x^2)) + (m1 <-> m2, conjugated), weighted by
    f(x) / (1 - x^2)^1.5."""
    avg, mag = (hi + lo) / 2, (hi - lo) / 2
    th = np.pi * (np.arange(Mq) + 0.5) / Mq
    x, s1 = np.cos(th), np.sin(th)
    keep = 1.0 - x * x >= 1e-4
    th, x, s1 = th[keep], x[keep], s1[keep]
    f = np.array([oracle.fermi_density(xi * mag + avg, kT, mu) for xi in x]) / s1 ** 3
    m = np.arange(M)
    T, S = np.cos(np.outer(th, m)), np.sin(np.outer(th, m))
    Q1 = (T - 1j * S) * (x[:, None] + 1j * m[None, :] * s1[:, None])
    Q2 = (T + 1j * S) * (x[:, None] - 1j * m[None, :] * s1[:, None])
    out = (T * f[:, None]).T @ Q1 + (Q2 * f[:, None]).T @ T
    w = np.where(m == 0, 1.0, 2.0) * kernel
    return out * (2 * np.pi * np.outer(w, w) / (Mq * mag * mag))


def _direct_optical(oracle, M, Mq, kT, mu, omega, lo, hi, kernel):
    """Optical coefficients as the node sum of the branch without FFTW
    (src/fastkpm.cpp:203-222): sum over nodes of T_m1(x + omega) T_m2(x)
    (f(x) - f(x + omega)) / (omega sqrt(1 - (x + omega)^2))."""
    avg, mag = (hi + lo) / 2, (hi - lo) / 2
    ws = omega / mag
    i0 = int(np.ceil(np.arccos(1.0 - ws) / np.pi * Mq - 0.5))
    x = np.cos(np.pi * (np.arange(i0, Mq) + 0.5) / Mq)
    xw = x + ws
    keep = (x >= -1 + 1e-4) & (xw <= 1 - 1e-4)
    x, xw = x[keep], xw[keep]
    e = x * mag + avg
    df = np.array([oracle.fermi_density(a, kT, mu) - oracle.fermi_density(a + omega, kT, mu)
                   for a in e])
    f = df / (ws * np.sqrt(1 - xw * xw))
    m = np.arange(M)
    T1, T2 = np.cos(np.outer(np.arccos(xw), m)), np.cos(np.outer(np.arccos(x), m))
    w = np.where(m == 0, 1.0, 2.0) * kernel
    return (T1 * f[:, None]).T @ T2 * (2 * np.pi * np.outer(w, w) / (Mq * mag * mag))


@pytest.mark.parametrize("M,Mq", [(96, 300), (64, 256), (50, 101)])
def test_conductivity_static_matches_node_sum(host, oracle, M, Mq):
    """The transform route (power-of-two and chirp-z lengths) against the
    direct node sum, a different algebraic route."""
    lo, hi, kT, mu = -3.2, 2.6, 0.05, -0.4
    g = host.jackson_kernel(M)
    got = host.conductivity_coefficients(M, Mq, kT, mu, 0.0, lo, hi)
    want = _direct_static(oracle, M, Mq, kT, mu, lo, hi, g)
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("M,Mq,omega", [(64, 300, 0.5), (40, 128, 1.3)])
def test_conductivity_optical_matches_node_sum(host, oracle, M, Mq, omega):
    lo, hi, kT, mu = -3.0, 3.0, 0.1, 0.2
    g = host.jackson_kernel(M)
    got = host.conductivity_coefficients(M, Mq, kT, mu, omega, lo, hi)
    want = _direct_optical(oracle, M, Mq, kT, mu, omega, lo, hi, g)
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()


def test_conductivity_large_M_is_fast_and_hermitian(host):
    """M = 2000 with Mq = 4000 (the expansion orders of the large workloads)
    finishes in seconds; the static coefficients satisfy
    C_{m2,m1} = conj(C_{m1,m2}) (the integrand's two terms swap under
    m1 <-> m2 with conjugation)."""
    import time
    M, Mq = 2000, 4000
    t0 = time.perf_counter()
    C = host.conductivity_coefficients(M, Mq, 0.05, 0.1, 0.0, -7.0, 7.0)
    assert time.perf_counter() - t0 < 10.0
    assert np.abs(C - C.conj().T).max() < 1e-10 * np.abs(C).max()


def test_moment_transform2_matches_direct_sum(host):
    """gamma[i][j] = sum T_m1(x_i) T_m2(x_j) (2 - delta)(2 - delta) g g mu[m1][m2]
    (the quadruple loop of src/fastkpm.cpp:258-289), at a chirp-z length."""
    rng = np.random.default_rng(9)
    M, Mq = 17, 45
    mu = rng.normal(size=(M, M)) + 1j * rng.normal(size=(M, M))
    g = host.jackson_kernel(M)
    w = np.where(np.arange(M) == 0, 1.0, 2.0) * g
    T = np.cos(np.outer(np.arange(M), np.pi * (np.arange(Mq) + 0.5) / Mq))
    want = T.T @ (np.outer(w, w) * mu) @ T
    got = host.moment_transform2(mu, Mq)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_moment_transform2_is_separable(host):
    """For moments a (x) b the 2-D transform is the outer product of the 1-D
    transforms (both apply the Jackson kernel)."""
    rng = np.random.default_rng(4)
    M, Mq = 14, 40
    a = rng.normal(size=M) + 1j * rng.normal(size=M)
    b = rng.normal(size=M) + 1j * rng.normal(size=M)
    got = host.moment_transform2(np.outer(a, b), Mq)
    ta = host.moment_transform(a.real, Mq) + 1j * host.moment_transform(a.imag, Mq)
    tb = host.moment_transform(b.real, Mq) + 1j * host.moment_transform(b.imag, Mq)
    want = np.outer(ta, tb)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
