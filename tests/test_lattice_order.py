"""The lattice row order (FKPM_OPT_ROW_ORDER) that set_H builds for a periodic
3-D lattice, computed on the host through fkpm_lattice_order: a permutation of
the block rows, one x-slab of n1/8 sites per eighth of the sweep, each walked
in y-tiles plane by plane along z; no order for 2-D lattices, thin lattices or
patterns with extra bonds."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import with_partner_bonds


def _coords(order, n1, n2):
    order = np.asarray(order)
    return order % n1, (order // n1) % n2, order // (n1 * n2)


@pytest.mark.parametrize("shape,b", [((64, 64, 64), 1), ((16, 8, 24), 2), ((24, 16, 16), 2),
                                     ((8, 8, 8), 8)])
def test_order_of_a_cubic_lattice(shape, b):
    from fastkpm_amd import lattice, lattice_order
    H = lattice.lattice(shape, b=b, seed=1, orbital_blocks=(b == 8))
    got = lattice_order(H)
    assert got is not None
    dims, order = got
    assert dims == shape
    nb = H.n_rows
    assert np.array_equal(np.sort(order), np.arange(nb))
    n1, n2, n3 = shape
    w = n1 // 8
    x, y, z = _coords(order, n1, n2)
    # XCD slab: positions [k nb/8, (k+1) nb/8) hold x in [k w, (k+1) w)
    assert np.array_equal(x // w, np.arange(nb) // (nb // 8))
    # inside a slab: the y-tile, then z, then y in the tile, then x
    ty = max(t for t in range(1, n2 + 1) if n2 % t == 0 and w * t <= 128)
    key = (((x // w) * (n2 // ty) + y // ty) * n3 + z) * ty * w + (y % ty) * w + x % w
    assert np.array_equal(key, np.arange(nb))
    # a row's +z neighbour is one tile-plane (ty w sites) on in the sweep,
    # except across the periodic boundary
    pos = np.empty(nb, dtype=np.int64)
    pos[order] = np.arange(nb)
    site = np.arange(nb)
    up = (site + n1 * n2) % nb
    inner = (site // (n1 * n2)) < n3 - 1
    assert np.all(pos[up[inner]] - pos[site[inner]] == ty * w)


def test_no_order_for_other_patterns():
    from fastkpm_amd import lattice, lattice_order
    assert lattice_order(lattice.lattice((256, 256), b=2, seed=1)) is None  # 2-D
    assert lattice_order(lattice.lattice((4, 64, 64), b=2, seed=1)) is None  # n1 < 8
    assert lattice_order(lattice.lattice((12, 12, 12), b=2, seed=1)) is None  # n1 % 8 != 0
    H = lattice.lattice((4, 4, 4), b=2, seed=1)
    assert lattice_order(H) is None  # fewer than 512 block rows
    shape = (16, 16, 16)
    H2 = with_partner_bonds(lattice.lattice(shape, b=2, seed=2), shape, 3)  # +-2 x bonds
    assert lattice_order(H2) is None  # a fourth bond distance


def test_bad_arguments():
    from fastkpm_amd import SpMatBsr, lattice_order
    from fastkpm_amd._capi import FkpmError
    bad = SpMatBsr(4, 4, 1, np.zeros(1, np.int32), np.array([7], np.int32),
                   np.array([0, 1, 1, 1, 1], np.int32), np.ones(1))
    with pytest.raises(FkpmError):
        lattice_order(bad)  # column index out of range
