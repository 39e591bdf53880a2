"""Batched engine (serving runtime) through the C-ABI: each request against
the CPU oracle (FP64 1e-10 relative, norm-wise), bitwise equality of a
request's results alone and in a mix (graph-captured or eager ticks,
admission-limited by an HBM budget), and the refusal paths."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import fermi_coeffs, rel_err, seeded_R, with_partner_bonds

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _req(oracle, kind, seed, M):
    """(H, es, c, R) of a small request; R dense from the oracle's set_R_correlated."""
    from fastkpm_amd import lattice, EnergyScale
    if kind == "b2_2d":
        shape, b, per = (6, 8), 2, (3, 4)
        H = lattice.lattice(shape, b=2, J=1.5, W=2.0, seed=seed)
    elif kind == "b2_3d":
        shape, b, per = (4, 4, 4), 2, (2, 2, 2)
        H = lattice.lattice(shape, b=2, J=1.0, W=4.0, seed=seed)
    else:
        shape, b, per = (3, 3, 2), 8, (3, 3, 2)
        H = lattice.lattice(shape, b=8, orbital_blocks=True, seed=seed)
    lam = np.linalg.eigvalsh(H.to_dense())
    es = EnergyScale(lam[0] - 0.2, lam[-1] + 0.2)
    c = fermi_coeffs(oracle, M, es.lo, es.hi, kT=0.05)
    groups = lattice.colouring(shape, per, b)
    R = seeded_R(oracle, groups, seed + 100)
    return H, es, c, R, groups


def _oracle_solve(oracle, H, es, c, R):
    orc = oracle.Engine()
    orc.set_H(H, es.lo, es.hi)
    orc.set_R(R)
    mu = orc.moments(len(c))
    D = orc.autodiff_matrix(c, H.like())
    return mu, D


MIX = [("b2_2d", 3, 40), ("b8", 5, 24), ("b2_3d", 7, 64), ("b2_2d", 9, 2), ("b8", 11, 36),
       ("b2_3d", 13, 30)]


@pytest.fixture(scope="module")
def mix(oracle):
    return [_req(oracle, k, s, M) for k, s, M in MIX]


def test_batch_fp64_matches_oracle(oracle, mix):
    from fastkpm_amd.serving import BatchEngine
    eng = BatchEngine("fp64", hbm_bytes=1 << 30, max_slots=8)
    ids = [eng.submit(H, es, c, R=R) for H, es, c, R, _ in mix]
    for (H, es, c, R, _), rid in zip(mix, ids):
        r = eng.wait(rid)
        mu_o, D_o = _oracle_solve(oracle, H, es, c, R)
        assert rel_err(r.mu, mu_o) < FP64_TOL
        assert abs(r.energy - float(np.dot(c, mu_o))) <= FP64_TOL * abs(np.dot(c, mu_o)) + 1e-12
        assert rel_err(r.D, D_o) < FP64_TOL
    st = eng.stats()
    assert st["captures"] >= 1 and st["ticks"] >= max(len(c) for _, _, c, _, _ in mix) + 2
    eng.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_alone_equals_mix_bitwise(mix, prec):
    """A request's results do not depend on what else is in flight, nor on
    graph capture, nor on the tick run length."""
    from fastkpm_amd.serving import BatchEngine
    alone = []
    eng = BatchEngine(prec, hbm_bytes=1 << 30, max_slots=8, graph=False)
    for H, es, c, R, _ in mix:
        alone.append(eng.wait(eng.submit(H, es, c, R=R)))
    eng.close()
    for graph, ticks in ((True, 16), (True, 5), (False, 3)):
        eng = BatchEngine(prec, hbm_bytes=1 << 30, max_slots=4, graph=graph, max_ticks=ticks)
        ids = [eng.submit(H, es, c, R=R) for H, es, c, R, _ in mix]
        for a, rid in zip(alone, ids):
            r = eng.wait(rid)
            assert np.array_equal(r.mu, a.mu), (graph, ticks)
            assert np.array_equal(r.D, a.D), (graph, ticks)
            assert r.energy == a.energy
        eng.close()


def test_batch_fp32_mixed_against_oracle(oracle, mix):
    from fastkpm_amd.serving import BatchEngine
    eng = BatchEngine("fp32", hbm_bytes=1 << 30, max_slots=8)
    ids = [eng.submit(H, es, c, R=R) for H, es, c, R, _ in mix]
    for (H, es, c, R, _), rid in zip(mix, ids):
        r = eng.wait(rid)
        mu_o, D_o = _oracle_solve(oracle, H, es, c, R)
        e_o = float(np.dot(c, mu_o))
        assert abs(r.energy - e_o) / abs(e_o) < 1e-6
        assert rel_err(r.D, D_o) < 1e-4
    eng.close()


def test_batch_admission_by_footprint(oracle, mix):
    """With an arena that holds about two requests, all still complete, each
    exactly as with room for all."""
    from fastkpm_amd.serving import BatchEngine, footprint
    fps = [footprint("fp64", H, R.shape[1], len(c)) for H, es, c, R, _ in mix]
    budget = int(sorted(fps)[-1] + sorted(fps)[-2] + (1 << 20))
    big = BatchEngine("fp64", hbm_bytes=1 << 30, max_slots=8)
    ref = [big.wait(big.submit(H, es, c, R=R)) for H, es, c, R, _ in mix]
    big.close()
    eng = BatchEngine("fp64", hbm_bytes=budget, max_slots=8, max_ticks=7)
    ids = [eng.submit(H, es, c, R=R) for H, es, c, R, _ in mix]
    for a, rid in zip(ref, ids):
        r = eng.wait(rid)
        assert np.array_equal(r.mu, a.mu) and np.array_equal(r.D, a.D)
    st = eng.stats()
    assert st["peak_bytes"] <= budget
    assert st["peak_slots"] < len(mix)
    eng.close()


def test_batch_probing_form_equals_dense(oracle):
    """Compact probing vectors give the same results as the dense R they stand for."""
    from fastkpm_amd import host
    from fastkpm_amd.serving import BatchEngine
    H, es, c, _, groups = _req(oracle, "b2_2d", 21, 48)
    R = host.set_R_correlated("cx_double", groups, host.RNG(77))
    probing = host.set_R_probing(groups, host.RNG(77))
    eng = BatchEngine("fp64", hbm_bytes=1 << 28, max_slots=4)
    a = eng.wait(eng.submit(H, es, c, R=R))
    b = eng.wait(eng.submit(H, es, c, probing=probing))
    assert np.array_equal(a.mu, b.mu) and np.array_equal(a.D, b.D)
    eng.close()


@pytest.mark.parametrize("shape,b,per", [((64, 64), 2, (2, 2)), ((5, 5, 4), 8, (1, 1, 2)),
                                         ((9, 7), 2, (3, 7))])
def test_batch_seeded_equals_host_probing(shape, b, per):
    """Probing vectors drawn on the device from the column seeds equal the
    host's set_R_correlated bit for bit: the results match exactly.  The
    first case gives every colour 1024 rows (generator blocks of 624 words
    are crossed inside a 64-row chunk)."""
    from fastkpm_amd import host, lattice
    from fastkpm_amd.serving import BatchEngine, gershgorin
    if b == 2:
        H = lattice.lattice(shape, b=2, J=1.2, W=1.0, seed=31)
    else:
        H = lattice.lattice(shape, b=8, orbital_blocks=True, seed=31)
    es = gershgorin(H)
    M = 24
    c = host.expansion_coefficients(M, 2 * M, host.FERMI_ENERGY, 0.05, 0.0, es.lo, es.hi)
    groups = lattice.colouring(shape, per, b)
    s = int(groups.max()) + 1
    probing = host.set_R_probing(groups, host.RNG(2024))
    seeds = host.master_seeds(host.RNG(2024), s)
    eng = BatchEngine("fp64", hbm_bytes=1 << 28, max_slots=4)
    a = eng.wait(eng.submit(H, es, c, probing=probing))
    d = eng.wait(eng.submit(H, es, c, seeded=(groups, seeds)))
    assert np.array_equal(a.mu, d.mu) and np.array_equal(a.D, d.D)
    eng.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_lattice_row_order(oracle, prec):
    """A 3-D lattice request with 4096 block rows takes the lattice row order
    (its slot sweeps the order's positions one range per XCD), in one class
    launch with a 2-D request that keeps row order: both against the oracle,
    and the ordered request bitwise the same alone."""
    from fastkpm_amd import host, lattice, lattice_order
    from fastkpm_amd.serving import BatchEngine, gershgorin
    reqs = []
    for shape, per, seed in (((16, 16, 16), (2, 2, 2), 51), ((64, 64), (2, 4), 52)):
        H = lattice.lattice(shape, b=2, J=1.0, W=2.0, seed=seed)
        es = gershgorin(H)
        M = 20
        c = fermi_coeffs(oracle, M, es.lo, es.hi, kT=0.05)
        R = seeded_R(oracle, lattice.colouring(shape, per, 2), seed + 100)
        reqs.append((H, es, c, R))
    assert lattice_order(reqs[0][0]) is not None and lattice_order(reqs[1][0]) is None
    assert reqs[0][3].shape[1] == reqs[1][3].shape[1] == 16  # the same class
    tol_mu, tol_d = (FP64_TOL, FP64_TOL) if prec == "fp64" else (2e-5, 1e-4)
    eng = BatchEngine(prec, hbm_bytes=1 << 30, max_slots=4)
    ids = [eng.submit(H, es, c, R=R) for H, es, c, R in reqs]
    res = [eng.wait(rid) for rid in ids]
    for (H, es, c, R), r in zip(reqs, res):
        mu_o, D_o = _oracle_solve(oracle, H, es, c, R)
        assert rel_err(r.mu, mu_o) < tol_mu
        assert rel_err(r.D, D_o) < tol_d
    H, es, c, R = reqs[0]
    alone = eng.wait(eng.submit(H, es, c, R=R))
    assert np.array_equal(alone.mu, res[0].mu) and np.array_equal(alone.D, res[0].D)
    eng.close()


def _dense_rows_b8(nb=20, extra=6, seed=3):
    """Hermitian 8x8-block matrix with rows of up to 2*extra+1 blocks: more
    than the row-resident matrix-core kernels take."""
    from fastkpm_amd import lattice
    rng = np.random.default_rng(seed)
    rows, cols, blocks = [], [], []
    for i in range(nb):
        x = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
        rows.append(i)
        cols.append(i)
        blocks.append((x + x.conj().T) / 2)
        for j in rng.choice([k for k in range(nb) if k != i], size=extra, replace=False):
            t = 0.5 * (rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))
            rows += [i, int(j)]
            cols += [int(j), i]
            blocks += [t, t.conj().T]
    return lattice.coo_to_bsr(nb, 8, rows, cols, np.array(blocks))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_b8_row_resident_and_general(oracle, prec):
    """8x8-block requests take a row-resident matrix-core class when every
    row has at most 8 blocks (on 7 block slots when no row has more than 7,
    else on 8) and the general one otherwise; all three in flight together,
    each against the oracle."""
    from fastkpm_amd import EnergyScale, lattice
    from fastkpm_amd.serving import BatchEngine
    Hd = _dense_rows_b8()
    assert int(np.max(np.diff(Hd.row_ptr))) > 8
    lam = np.linalg.eigvalsh(Hd.to_dense())
    esd = EnergyScale(lam[0] - 0.2, lam[-1] + 0.2)
    groups = (np.arange(8)[None, :] + 8 * (np.arange(Hd.n_rows) % 3)[:, None]).reshape(-1)
    Rd = seeded_R(oracle, groups.astype(np.int32), 41)
    Md = 28
    cd = fermi_coeffs(oracle, Md, esd.lo, esd.hi, kT=0.05)
    Hl, esl, cl, Rl, _ = _req(oracle, "b8", 43, 36)
    shape8 = (4, 4, 4)
    H8 = with_partner_bonds(lattice.lattice(shape8, b=8, orbital_blocks=True, seed=45), shape8, 46)
    assert int(np.min(np.diff(H8.row_ptr))) == 8 and int(np.max(np.diff(H8.row_ptr))) == 8
    lam8 = np.linalg.eigvalsh(H8.to_dense())
    es8 = EnergyScale(lam8[0] - 0.2, lam8[-1] + 0.2)
    c8 = fermi_coeffs(oracle, 20, es8.lo, es8.hi, kT=0.05)
    R8 = seeded_R(oracle, lattice.colouring(shape8, (2, 2, 2), 8), 47)
    eng = BatchEngine(prec, hbm_bytes=1 << 30, max_slots=4)
    reqs = [(Hd, esd, cd, Rd), (Hl, esl, cl, Rl), (H8, es8, c8, R8)]
    ids = [eng.submit(H, es, c, R=R) for H, es, c, R in reqs]
    tol_mu, tol_d = (FP64_TOL, FP64_TOL) if prec == "fp64" else (2e-5, 1e-4)
    for (H, es, c, R), rid in zip(reqs, ids):
        r = eng.wait(rid)
        mu_o, D_o = _oracle_solve(oracle, H, es, c, R)
        assert rel_err(r.mu, mu_o) < tol_mu
        assert rel_err(r.D, D_o) < tol_d
    eng.close()


def test_batch_refusals(oracle):
    from fastkpm_amd import lattice, EnergyScale
    from fastkpm_amd._capi import FkpmError
    from fastkpm_amd.serving import BatchEngine
    eng = BatchEngine("fp64", hbm_bytes=1 << 20, max_slots=2)
    H4 = lattice.lattice((4, 4), b=4, orbital_blocks=True, seed=1)
    R4 = np.ones((H4.n_rows * 4, 2), dtype=np.complex128)
    with pytest.raises(FkpmError, match="UNSUP"):
        eng.submit(H4, EnergyScale(-10, 10), np.ones(8), R=R4)
    Hb = lattice.lattice((16, 16, 16), b=2, seed=2)  # 8192 orbitals: 2 MiB per vector
    Rb = np.ones((Hb.n_rows * 2, 16), dtype=np.complex128)
    with pytest.raises(FkpmError, match="budget"):
        eng.submit(Hb, EnergyScale(-12, 12), np.ones(64), R=Rb)
    # a missing diagonal block
    P = lattice.coo_to_bsr(2, 2, [0, 1], [1, 0], np.ones((2, 2, 2)))
    with pytest.raises(FkpmError, match="diagonal"):
        eng.submit(P, EnergyScale(-3, 3), np.ones(4), R=np.ones((4, 1), dtype=np.complex128))
    # seeded: a colour outside [0, s), and as many colours as rows
    H2 = lattice.lattice((4, 4), b=2, seed=3)
    g = lattice.colouring((4, 4), (2, 2), 2)
    with pytest.raises(FkpmError, match="colour"):
        eng.submit(H2, EnergyScale(-6, 6), np.ones(4), seeded=(g, np.arange(3, dtype=np.uint32)))
    with pytest.raises(FkpmError, match="colours"):
        eng.submit(H2, EnergyScale(-6, 6), np.ones(4),
                   seeded=(np.arange(32), np.arange(32, dtype=np.uint32)))
    eng.close()


def test_batch_close_wakes_waiters():
    """Closing the engine while other threads wait on a running and on a queued
    request: each waiter returns (the result, or FKPM_E_STATE for a request
    the engine did not finish) and close returns after them."""
    import threading
    from fastkpm_amd import serving
    from fastkpm_amd._capi import FkpmError
    big = serving.RequestSpec(3, 2, 40, 4, 1024, 2.0, 17)  # N = 128000, S = 128, M = 1024
    H, es, c, sd = serving.build_request(big)
    fp = serving.footprint("fp64", H, len(sd[1]), len(c))
    # room for one of them at a time: the second stays queued
    eng = serving.BatchEngine("fp64", hbm_bytes=int(fp * 1.5), max_slots=4)
    ids = [eng.submit(H, es, c, seeded=sd) for _ in range(2)]
    out = {}

    def waiter(k, rid):
        try:
            out[k] = eng.wait(rid, want_D=False).energy
        except FkpmError as e:
            out[k] = e

    threads = [threading.Thread(target=waiter, args=(k, rid)) for k, rid in enumerate(ids)]
    for t in threads:
        t.start()
    import time
    time.sleep(0.3)  # both waiters inside fkpm_batch_wait, the first request running
    eng.close()
    for t in threads:
        t.join(timeout=60)
        assert not t.is_alive()
    assert len(out) == 2
    # the queued request cannot have finished: the arena held one at a time
    assert any(isinstance(v, FkpmError) and v.code == -2 for v in out.values()), out
