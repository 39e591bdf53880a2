"""Host side of the serving stream (no GPU): request generation, spectral
bounds and compact probing vectors."""
from __future__ import annotations

import numpy as np

from fastkpm_amd import host, lattice
from fastkpm_amd.serving import build_request, draw_spec, gershgorin


def test_gershgorin_bounds_contain_the_spectrum():
    for H in (lattice.lattice((6, 6), b=2, J=2.0, W=3.0, seed=4),
              lattice.lattice((3, 3, 2), b=8, orbital_blocks=True, seed=5)):
        lam = np.linalg.eigvalsh(H.to_dense())
        es = gershgorin(H)
        assert es.lo < lam[0] and lam[-1] < es.hi


def test_probing_form_is_the_dense_R():
    for shape, per, b in (((8, 8), (4, 4), 2), ((4, 4, 2), (2, 2, 2), 8)):
        g = lattice.colouring(shape, per, b)
        R = host.set_R_correlated("cx_double", g, host.RNG(19))
        rcol, rval, s = host.set_R_probing(g, host.RNG(19))
        D = np.zeros_like(R)
        D[np.arange(len(g)), rcol] = rval
        assert s == R.shape[1] and np.array_equal(D, R)
    # colour count reaching n: identity, no draws
    rcol, rval, s = host.set_R_probing(np.arange(6), host.RNG(1))
    assert s == 6 and np.array_equal(rcol, np.arange(6)) and np.all(rval == 1)


def test_stream_specs_stay_in_range():
    rng = np.random.default_rng(0)
    for _ in range(200):
        sp = draw_spec(rng)
        assert 4096 <= sp.N <= 1 << 20
        assert 256 <= sp.M <= 2048 and sp.M % 2 == 0
        assert 16 <= sp.S <= 128
        assert 0.0 <= sp.W <= 6.0


def test_build_request_small():
    rng = np.random.default_rng(3)
    sp = draw_spec(rng, n_max=5000, m_max=256)
    H, es, c, (groups, seeds) = build_request(sp)
    assert H.n_rows * H.b_len == sp.N and len(c) == sp.M and len(seeds) == sp.S
    assert es.lo < es.hi and len(groups) == sp.N
    assert groups.min() == 0 and groups.max() == sp.S - 1


def test_master_seeds_are_the_column_seeds():
    """set_R_probing seeds column j's generator with the j-th master draw:
    the seeds of the device-side generation are those draws, and the master
    engine ends in the same state either way."""
    from fastkpm_amd import host, lattice
    groups = lattice.colouring((6, 4), (3, 2), 2)
    s = int(groups.max()) + 1
    a, b = host.RNG(41), host.RNG(41)
    seeds = host.master_seeds(a, s)
    rcol, rval, s2 = host.set_R_probing(groups, b)
    assert s2 == s and a() == b()
    # column 0's phases from mt19937(seeds[0]): uniform_int(0, 3) is the top two bits
    g = host.RNG(int(seeds[0]))
    rows = np.flatnonzero(groups == 0)
    phases = np.array([1, 1j, -1, -1j])[[g() >> 30 for _ in rows]]
    assert np.array_equal(rval[rows], phases) and np.all(rcol[rows] == 0)
