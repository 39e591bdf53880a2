"""Seeded synthetic Hamiltonians in BSR form, and probing colourings.

These are the workloads the engine is measured on: periodic square / cubic
lattices with nearest-neighbour hopping, optional Kondo coupling -J s_i.sigma
to random classical spins, Anderson on-site disorder, or dense random b x b
orbital blocks.  Every lattice stores an explicit diagonal block per site.
"""
from __future__ import annotations

import numpy as np

from .engine import SpMatBsr


def coo_to_bsr(nb: int, b: int, rows, cols, blocks, dtype=np.complex128) -> SpMatBsr:
    """COO blocks -> sorted BSR with duplicate (i, j) blocks summed, the
    semantics of SpMatBsr::build (src/spmat.cpp:46-129).  blocks[k] is a b x b
    array (row index bi, column index bj)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    blocks = np.asarray(blocks).reshape(len(rows), b, b)
    order = np.lexsort((cols, rows))
    rows, cols, blocks = rows[order], cols[order], blocks[order]
    key = rows * nb + cols
    uniq, first = np.unique(key, return_index=True)
    summed = np.add.reduceat(blocks, first, axis=0) if len(rows) else blocks
    r = (uniq // nb).astype(np.int32)
    c = (uniq % nb).astype(np.int32)
    row_ptr = np.zeros(nb + 1, dtype=np.int32)
    np.add.at(row_ptr, r.astype(np.int64) + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32)
    # column-major blocks: element (bi, bj) at bj*b + bi
    val = np.ascontiguousarray(np.transpose(summed, (0, 2, 1))).reshape(-1).astype(dtype)
    return SpMatBsr(nb, nb, b, r, c, row_ptr, val)


def _neighbours(shape):
    """Site index and +1 neighbour along each axis of a periodic lattice."""
    dims = len(shape)
    idx = np.arange(int(np.prod(shape))).reshape(shape, order="F")
    out = []
    for d in range(dims):
        out.append((idx.reshape(-1, order="F"), np.roll(idx, -1, axis=d).reshape(-1, order="F")))
    return out


def _coords(shape):
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return [g.reshape(-1, order="F") for g in grids]


def _random_spins(rng, n):
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def lattice(shape, b: int = 1, t: float = -1.0, J: float = 0.0, W: float = 0.0,
            orbital_blocks: bool = False, seed: int = 0, dtype=np.complex128) -> SpMatBsr:
    """Periodic hypercubic lattice of the given shape (site index x + Lx*y + ...).

    b = 1: scalar hopping t, on-site disorder U[-W/2, W/2].
    b = 2: spin-1/2 orbitals, hopping t*I2, on-site eps_i*I2 - J s_i.sigma with
           s_i uniform on the sphere.
    orbital_blocks: dense b x b blocks, on-site random Hermitian (X + X^+)/2 with
           complex Gaussian(0,1) X, bonds complex Gaussian(0, 0.5) T with the
           Hermitian partner T^+.
    """
    rng = np.random.default_rng(seed)
    nsite = int(np.prod(shape))
    rows, cols, blks = [], [], []
    # on-site
    if orbital_blocks:
        X = rng.standard_normal((nsite, b, b)) + 1j * rng.standard_normal((nsite, b, b))
        onsite = (X + np.conj(np.transpose(X, (0, 2, 1)))) / 2
    else:
        onsite = np.zeros((nsite, b, b), dtype=np.complex128)
        eps = rng.uniform(-W / 2, W / 2, nsite) if W else np.zeros(nsite)
        onsite += eps[:, None, None] * np.eye(b)[None]
        if b == 2 and J:
            s = _random_spins(rng, nsite)
            sx, sy, sz = s[:, 0], s[:, 1], s[:, 2]
            onsite[:, 0, 0] += -J * sz
            onsite[:, 1, 1] += J * sz
            onsite[:, 0, 1] += -J * (sx - 1j * sy)
            onsite[:, 1, 0] += -J * (sx + 1j * sy)
    rows.append(np.arange(nsite))
    cols.append(np.arange(nsite))
    blks.append(onsite)
    for src, dst in _neighbours(shape):
        if orbital_blocks:
            T = (rng.standard_normal((nsite, b, b)) + 1j * rng.standard_normal((nsite, b, b))) \
                * np.sqrt(0.5)
        else:
            T = np.broadcast_to(t * np.eye(b, dtype=np.complex128), (nsite, b, b))
        Th = np.conj(np.transpose(T, (0, 2, 1)))
        rows += [src, dst]
        cols += [dst, src]
        blks += [T, Th]
    H = coo_to_bsr(nsite, b, np.concatenate(rows), np.concatenate(cols),
                   np.concatenate(blks, axis=0), dtype=np.complex128)
    if not np.issubdtype(np.dtype(dtype), np.complexfloating):
        if np.abs(H.val.imag).max(initial=0.0) != 0.0:
            raise ValueError("a real dtype needs a real Hamiltonian")
        H.val = H.val.real.astype(dtype)
    else:
        H.val = H.val.astype(dtype)
    return H


def current(H: SpMatBsr, shape, axis: int, real: bool = False) -> SpMatBsr:
    """Current operator j = i [H, X_axis] of a periodic lattice Hamiltonian
    (block rows = sites, index x + Lx*y + ...): block (i, j) is
    i (x_j - x_i) H_ij with the minimum-image displacement; blocks with no
    displacement (on-site, bonds along other axes) are left out of the
    pattern.  real=True drops the factor i (a real antisymmetric operator of
    the same pattern, for real engines)."""
    xs = _coords(shape)[axis].astype(np.int64)
    L = shape[axis]
    b = H.b_len
    rows = np.repeat(np.arange(H.n_rows), np.diff(np.asarray(H.row_ptr)))
    cols = np.asarray(H.col_idx, dtype=np.int64)
    d = (xs[cols] - xs[rows] + L // 2) % L - L // 2
    keep = d != 0
    # column-major blocks: element (bi, bj) at bj*b + bi
    blk = np.transpose(np.asarray(H.val).reshape(-1, b, b), (0, 2, 1))[keep]
    fac = d[keep].astype(np.float64) if real else 1j * d[keep]
    out = coo_to_bsr(H.n_rows, b, rows[keep], cols[keep], fac[:, None, None] * blk)
    out.val = out.val.real.astype(H.val.dtype) if real else out.val.astype(H.val.dtype)
    return out


def colouring(shape, periods, b: int = 1) -> np.ndarray:
    """Probing colours: orbital + b * spatial colour, the spatial colour being
    the site position modulo `periods` (e.g. (4, 4) on a square lattice)."""
    xs = _coords(shape)
    col = np.zeros(int(np.prod(shape)), dtype=np.int64)
    mult = 1
    for x, p in zip(xs, periods):
        col += (x % p) * mult
        mult *= p
    groups = (np.arange(b)[None, :] + b * col[:, None]).reshape(-1)
    return groups.astype(np.int32)


# The measured configurations: lattice arguments, probing periods, M, S.
CONFIGS = {
    "C1": dict(shape=(32, 32), b=1, t=-1.0, periods=(4, 4), M=500, kT=0.0, mu=0.0,
               es=(-4.1, 4.1), real=True),
    "C2": dict(shape=(256, 256), b=2, t=-1.0, J=2.0, periods=(4, 8), M=1000, kT=0.01, mu=0.0),
    "C3": dict(shape=(64, 64, 64), b=2, t=-1.0, J=1.0, W=4.0, periods=(4, 4, 4), M=2000,
               kT=0.01, mu=0.0),
    "C4": dict(shape=(32, 32, 32), b=8, orbital_blocks=True, periods=(4, 4, 2), M=1024,
               kT=0.01, mu=0.0),
}


def make_config(name: str, seed: int = 0, scale_shape=None, dtype=np.complex128):
    """Hamiltonian and colouring of a configuration (optionally on a smaller
    lattice of the same kind, for parity tests)."""
    cfg = dict(CONFIGS[name])
    shape = tuple(scale_shape) if scale_shape else cfg["shape"]
    H = lattice(shape, b=cfg["b"], t=cfg.get("t", -1.0), J=cfg.get("J", 0.0), W=cfg.get("W", 0.0),
                orbital_blocks=cfg.get("orbital_blocks", False), seed=seed, dtype=dtype)
    groups = colouring(shape, cfg["periods"], cfg["b"])
    return H, groups, cfg
