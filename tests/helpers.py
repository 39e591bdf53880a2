"""Shared helpers of the parity tests: seeded inputs and error measures."""
from __future__ import annotations

import numpy as np


def rel_err(a, b) -> float:
    """max |a - b| / max |b| (norm-wise relative error)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b)) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0)) if b.size else 0.0


def seeded_R(oracle, groups, seed, is_complex=True):
    """R from Engine::set_R_correlated(groups, rng) with rng = mt19937(seed)."""
    rng = oracle.MT19937(seed)
    return oracle.set_R_correlated(groups, rng, is_complex=is_complex)


def fermi_coeffs(oracle, M, lo, hi, kT=0.01, mu=0.0, kind=0):
    """expansion_coefficients(M, 2M, fermi_energy(., kT, mu), es)."""
    return oracle.expansion_coefficients_fermi(M, 2 * M, kind, kT, mu, lo, hi)
