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
