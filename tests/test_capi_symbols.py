"""The native libraries load and export every entry point their headers
declare (no compute calls: these run without a GPU)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fkpm_\w+)\s*\(", text)))


def test_device_library_exports_every_declared_symbol():
    from fastkpm_amd import _capi
    L = ctypes.CDLL(str(_capi.LIB_PATH))
    names = _declared("fastkpm_hip.h")
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares a signature for every one of them
    assert set(names) <= set(_capi.SIGNATURES), set(names) - set(_capi.SIGNATURES)


def test_host_library_exports_every_declared_symbol():
    from fastkpm_amd import host
    L = host.lib()
    names = _declared("fastkpm_host.h")
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(host.SIGNATURES)


def test_open_without_device_fails_cleanly():
    from fastkpm_amd import _capi, device_count
    if device_count() > 0:
        pytest.skip("a HIP device is visible")
    h = ctypes.c_void_p()
    rc = _capi.lib().fkpm_dev_open(0, _capi.C64, _capi.PREC_FP64, ctypes.byref(h))
    assert rc == -4 and not h.value
    assert b"no HIP device" in _capi.lib().fkpm_last_error()


def test_bad_arguments_rejected_before_any_device_work():
    from fastkpm_amd import _capi
    h = ctypes.c_void_p()
    assert _capi.lib().fkpm_dev_open(0, 9, 0, ctypes.byref(h)) == -1
    assert _capi.lib().fkpm_dev_open(0, 3, 7, ctypes.byref(h)) == -1
    assert _capi.lib().fkpm_moments(None, 10, None) == -1
