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


def with_partner_bonds(H1, shape, seed):
    """H1 plus random Hermitian b x b blocks on the (+2, 0, 0) bond of a lattice
    4 sites wide in x (+2 and -2 are the same site): one more block per row
    (8 on a cubic lattice)."""
    from fastkpm_amd import lattice
    rng = np.random.default_rng(seed)
    b = H1.b_len
    nsite = H1.n_rows
    idx = np.arange(nsite).reshape(shape, order="F")
    src = idx.reshape(-1, order="F")
    dst = np.roll(idx, -2, axis=0).reshape(-1, order="F")
    T = (rng.standard_normal((nsite, b, b)) + 1j * rng.standard_normal((nsite, b, b))) * 0.3
    rows = [H1.row_idx, src, dst]
    cols = [H1.col_idx, dst, src]
    blks = [H1.val.reshape(-1, b, b).transpose(0, 2, 1), T, np.conj(T.transpose(0, 2, 1))]
    return lattice.coo_to_bsr(nsite, b, np.concatenate(rows), np.concatenate(cols),
                              np.concatenate(blks, axis=0))
