"""The C++ fkpm:: API (include/fastkpm.h) driven by an application-style
program, tests/cpp/engine_flow.cpp: host checks on the CPU, and on the GPU the
full flow compared with the oracle on the H and R the program used."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from helpers import rel_err

BIN = Path(__file__).resolve().parent / "cpp" / "engine_flow"


def _bin():
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    return str(BIN)


def test_cpp_api_host_checks():
    r = subprocess.run([_bin(), "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "host OK" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("autotune", ["0", "1"])
def test_cpp_api_device_flow_matches_oracle(oracle, tmp_path, autotune):
    """autotune 1: FKPM_AUTOTUNE=1, the tuner runs at the first use of each
    shape inside the C++ engine."""
    import os
    env = dict(os.environ, FKPM_AUTOTUNE=autotune)
    r = subprocess.run([_bin(), "device", str(tmp_path)], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    from fastkpm_amd import SpMatBsr
    meta = (tmp_path / "meta.txt").read_text().split()
    n, s, M, nb, nnzb = (int(v) for v in meta[:5])
    lo, hi, E = (float(v) for v in meta[5:8])
    M2 = int(meta[8])
    sigma = complex(float(meta[9]), float(meta[10]))
    rd = lambda name, dt: np.fromfile(tmp_path / name, dtype=dt)
    H = SpMatBsr(nb, nb, 2, rd("H_row_idx.bin", np.int32), rd("H_col_idx.bin", np.int32),
                 rd("H_row_ptr.bin", np.int32), rd("H_val.bin", np.complex128))
    assert H.n_blocks() == nnzb
    R = rd("R.bin", np.complex128).reshape(s, n).T
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    mu_o = orc.moments(M)
    assert rel_err(rd("mu.bin", np.float64), mu_o) < 1e-10
    cE = rd("cE.bin", np.float64)
    assert E == pytest.approx(float(np.dot(cE, mu_o)), rel=1e-10)
    d_o = orc.autodiff_matrix(cE, H)
    assert rel_err(rd("D_autodiff.bin", np.complex128), d_o) < 1e-10
    s_o = orc.stoch_matrix(rd("cF.bin", np.float64), H)
    assert rel_err(rd("D_stoch.bin", np.complex128), s_o) < 1e-10
    # the program drew R through the product's set_R_correlated; the oracle's
    # restatement of engine.cpp draws the same phases
    groups = np.argmax(np.abs(R) > 0, axis=1)
    assert np.array_equal(oracle.set_R_correlated(groups, oracle.MT19937(2024)), R)
    # two-current moments (x current, probe columns in chunks of 5) and the
    # static conductivity from them
    J = SpMatBsr(nb, nb, 2, rd("j_row_idx.bin", np.int32), rd("j_col_idx.bin", np.int32),
                 rd("j_row_ptr.bin", np.int32), rd("j_val.bin", np.complex128))
    assert J.n_blocks() == int(meta[11])
    mu2_o = orc.moments2_v1(M2, J, J)
    mu2 = rd("mu2.bin", np.complex128).reshape(M2, M2)
    assert rel_err(mu2, mu2_o) < 1e-10
    from fastkpm_amd import host
    c2 = host.conductivity_coefficients(M2, 2 * M2, 0.05, 0.0, 0.0, lo, hi)
    s_o = complex(np.sum(c2 * mu2_o))
    assert abs(sigma - s_o) < 1e-10 * abs(s_o)


ADAPTER = Path(__file__).resolve().parent / "cpp" / "ref_adapter"

# element type -> (numpy type of the dumped values, tolerance of the D outputs):
# the adapter opens the C-ABI in FP64 for every element type, so moments agree
# with the oracle to the FP64 bar; single-precision D values are rounded once
# to float on the way out (2^-24 relative per element).
ADAPTER_TYPES = {"cx_double": (np.complex128, 1e-10), "double": (np.float64, 1e-10),
                 "cx_float": (np.complex64, 1e-7), "float": (np.float32, 1e-7)}


@pytest.mark.gpu
@pytest.mark.parametrize("etype", list(ADAPTER_TYPES))
def test_reference_side_adapter_matches_oracle(oracle, tmp_path, etype):
    """tests/cpp/ref_adapter.cpp: an fkpm::Engine<T> forwarding to the C-ABI the
    way a maintainer of the reference would bind it (block-row counts passed as
    SpMatBsr::n_rows), on a 12 x 12 two-orbital lattice (b = 2) with correlated
    probing R, against the oracle on the same H and R."""
    if not ADAPTER.exists():
        pytest.fail(f"{ADAPTER} is not built (run __graft_entry__.build())")
    r = subprocess.run([str(ADAPTER), etype, str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    from fastkpm_amd import SpMatBsr
    dt, tol = ADAPTER_TYPES[etype]
    meta = (tmp_path / "meta.txt").read_text().split()
    n, s, M, nb, nnzb = (int(v) for v in meta[:5])
    lo, hi = (float(v) for v in meta[5:7])
    rd = lambda name, t: np.fromfile(tmp_path / name, dtype=t)
    H = SpMatBsr(nb, nb, 2, rd("H_row_idx.bin", np.int32), rd("H_col_idx.bin", np.int32),
                 rd("H_row_ptr.bin", np.int32), rd("H_val.bin", dt))
    assert H.n_blocks() == nnzb and n == 2 * nb
    R = rd("R.bin", dt).reshape(s, n).T
    assert s == 8 and np.all(np.count_nonzero(R, axis=1) == 1)
    orc = oracle.Engine()
    orc.set_H(H, lo, hi)
    orc.set_R(R)
    mu_o = orc.moments(M)
    assert rel_err(rd("mu.bin", np.float64), mu_o) < 1e-10
    assert mu_o[0] == pytest.approx(n, rel=1e-12)
    d_o = orc.autodiff_matrix(rd("cE.bin", np.float64), H)
    assert rel_err(rd("D_autodiff.bin", dt), d_o) < tol
    s_o = orc.stoch_matrix(rd("cF.bin", np.float64), H)
    assert rel_err(rd("D_stoch.bin", dt), s_o) < tol
