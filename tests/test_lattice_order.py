"""The lattice row order (FKPM_OPT_ROW_ORDER) that set_H builds for a periodic
3-D lattice, computed on the host through fkpm_lattice_order: a permutation of
the block rows, one x-slab of about n1/8 sites per XCD, each walked in y-tiles
plane by plane along z; no order for 2-D lattices, thin lattices or patterns
with extra bonds."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import with_partner_bonds


def _coords(order, n1, n2):
    order = np.asarray(order)
    return order % n1, (order // n1) % n2, order // (n1 * n2)


@pytest.mark.parametrize("shape,b", [((64, 64, 64), 1), ((16, 8, 24), 2), ((24, 16, 16), 2),
                                     ((8, 8, 8), 8), ((12, 12, 12), 2), ((76, 10, 9), 2)])
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
    x, y, z = _coords(order, n1, n2)
    # XCD slabs: x in [floor(k n1 / 8), floor((k+1) n1 / 8)), in order of k
    edges = [k * n1 // 8 for k in range(9)]
    slab = np.searchsorted(edges, x, side="right") - 1
    assert np.all(np.diff(slab) >= 0) and slab[0] == 0 and slab[-1] == 7
    wmax = -(-n1 // 8)
    ty = max(t for t in range(1, n2 + 1) if n2 % t == 0 and wmax * t <= 128)
    # inside a slab: the y-tile, then z, then y in the tile, then x
    w = np.diff(edges)[slab]
    x0 = np.asarray(edges)[slab]
    key = ((y // ty) * n3 + z) * ty * w + (y % ty) * w + (x - x0)
    start = np.asarray([0] + list(np.cumsum(np.diff(edges) * n2 * n3)))[slab]
    assert np.array_equal(start + key, np.arange(nb))
    # a row's +z neighbour is one tile-plane (ty w sites) on in the sweep,
    # except across the periodic boundary
    pos = np.empty(nb, dtype=np.int64)
    pos[order] = np.arange(nb)
    site = np.arange(nb)
    up = (site + n1 * n2) % nb
    inner = (site // (n1 * n2)) < n3 - 1
    ws = np.diff(edges)[np.searchsorted(edges, site % n1, side="right") - 1]
    assert np.all(pos[up[inner]] - pos[site[inner]] == ty * ws[inner])


def test_no_order_for_other_patterns():
    from fastkpm_amd import lattice, lattice_order
    assert lattice_order(lattice.lattice((256, 256), b=2, seed=1)) is None  # 2-D
    assert lattice_order(lattice.lattice((4, 64, 64), b=2, seed=1)) is None  # n1 < 8
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
