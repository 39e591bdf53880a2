"""fastkpm_amd -- MI355X (gfx950) stochastic Chebyshev (KPM) engine.

The product is native: libfastkpm_hip.so (HIP kernels + the C-ABI of
include/fastkpm_hip.h) and libfastkpm.so (the C++ fkpm:: API of
include/fastkpm.h).  This package is the Python mirror of the engine interface
over the C-ABI, plus the synthetic lattice generators used by the tests and the
benchmark.
"""
from .engine import (EnergyScale, HipEngine, SpMatBsr, bandwidth_probe, device_count,  # noqa: F401
                     lattice_order)
from . import lattice  # noqa: F401
