"""Properties at the full sizes of the benchmark's workloads, where the oracle
would take minutes: C2 (256 x 256 sites x 2 spin orbitals, N = 131072, 64
probing colours), C3 (64^3 x 2, N = 524288, 128 colours) and C4 (32^3 sites
x 8 orbitals, N = 262144, 256 colours, matrix cores): relations the
mathematics fixes whatever the size, on a short expansion (M = 64) so each
test runs in seconds.

- probing R: mu_0 = sum |R|^2 = N;
- stoch_matrix: the diagonal of D = (R xi^+ + xi R^+) / 2 sums to
  Re <R, xi> = sum_m c_m mu_m, D is Hermitian, and the result from the kept
  forward vectors equals the Clenshaw recurrence's (C4, with no table: the
  Clenshaw result equals the forward order's);
- autodiff_matrix: sum_r D_rr = dE/d(eps) for H + eps I (central finite
  difference of E = sum_m c_m mu_m at fixed energy scale);
- FP32-mixed energy within 1e-6 of FP64.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

M = 64


@pytest.fixture(scope="module", params=["C2", "C3", "C4"])
def workload(request):
    from fastkpm_amd import HipEngine, lattice
    H, groups, cfg = lattice.make_config(request.param, seed=2024)
    eng = HipEngine("cx_double", "fp64")
    es = eng.energy_scale(H, 0.05)  # Lanczos bounds with 5 % slack
    eng.close()
    return H, groups, cfg, es


def _diag_sum(D) -> float:
    """sum over orbitals r of Re D_rr (diagonal elements of diagonal blocks)."""
    b = D.b_len
    blocks = D.val.reshape(-1, b * b)
    diag = np.nonzero(D.row_idx == D.col_idx)[0]
    return float(np.sum(blocks[diag][:, [bi * b + bi for bi in range(b)]].real))


def _engine(H, groups, es, prec="fp64", store=1):
    from fastkpm_amd import HipEngine, host
    eng = HipEngine("cx_double", prec)
    eng.set_store_alpha(store)
    eng.set_R_correlated(groups, host.RNG(7))
    eng.set_H(H, es)
    return eng


def test_fullsize_moments_and_stoch_trace(workload):
    from fastkpm_amd import host
    H, groups, cfg, es = workload
    N = H.n_rows * H.b_len
    eng = _engine(H, groups, es, store=2)
    mu = eng.moments(M)
    table = eng.alpha_table()  # not with the matrix-core autodiff step (C4)
    assert abs(mu[0] - N) <= 1e-12 * N, mu[0]
    c = host.expansion_coefficients(M, 2 * M, host.FERMI_DENSITY, cfg["kT"], cfg["mu"], es.lo,
                                    es.hi)
    D = H.like()
    eng.stoch_matrix(c, D)
    assert eng.stoch_method() == ("table" if table else "clenshaw")
    tr = _diag_sum(D)
    expect = float(np.dot(c, mu))
    assert abs(tr - expect) <= 1e-10 * abs(expect), (tr, expect)
    # Hermitian: block (i, j) = block (j, i)^+
    k = int(np.nonzero(D.row_idx != D.col_idx)[0][0])
    kt = D.find_index(int(D.col_idx[k]), int(D.row_idx[k]))
    assert np.allclose(D.block(k), D.block(kt).conj().T, rtol=0, atol=1e-14 * np.abs(D.val).max())
    # the same D from the other method (Clenshaw, or the forward order)
    eng.set_stoch_method("clenshaw" if table else "forward")
    D2 = H.like()
    eng.stoch_matrix(c, D2)
    assert rel_err(D2.val, D.val) < 1e-11
    eng.close()


def test_fullsize_autodiff_finite_difference(workload):
    from fastkpm_amd import host
    H, groups, cfg, es = workload
    c = host.expansion_coefficients(M, 2 * M, host.FERMI_ENERGY, cfg["kT"], cfg["mu"], es.lo,
                                    es.hi)
    eng = _engine(H, groups, es)
    eng.moments(M)
    D = H.like()
    eng.autodiff_matrix(c, D)
    grad = _diag_sum(D)
    # E(H + eps I) at the same energy scale, central difference
    eps = 1e-4
    b = H.b_len
    diag = np.nonzero(H.row_idx == H.col_idx)[0]
    idx = (diag[:, None] * b * b + np.array([bi * b + bi for bi in range(b)])[None, :]).ravel()
    energies = []
    for s in (1.0, -1.0):
        Hp = H.like(H.val.dtype)
        Hp.val[:] = H.val
        Hp.val[idx] += s * eps
        eng.set_H(Hp, es)
        energies.append(float(np.dot(c, eng.moments(M))))
    fd = (energies[0] - energies[1]) / (2 * eps)
    assert abs(fd - grad) <= 1e-6 * abs(grad), (fd, grad)
    eng.close()


def test_fullsize_fp32_energy(workload):
    from fastkpm_amd import host
    H, groups, cfg, es = workload
    c = host.expansion_coefficients(M, 2 * M, host.FERMI_ENERGY, cfg["kT"], cfg["mu"], es.lo,
                                    es.hi)
    e = []
    for prec in ("fp64", "fp32"):
        eng = _engine(H, groups, es, prec=prec)
        e.append(float(np.dot(c, eng.moments(M))))
        eng.close()
    assert abs(e[1] - e[0]) <= 1e-6 * abs(e[0]), e
