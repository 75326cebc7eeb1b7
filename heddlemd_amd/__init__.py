"""heddlemd_amd -- MI355X-native HeddleMD MD step.

The MD step (neighbour list, LJ + real-space Ewald pair forces, SPME
reciprocal space with an own LDS FFT, velocity Verlet, SETTLE, Nose-Hoover
chain, FP64 reductions) runs as hand-written HIP kernels for gfx950 inside
``libheddle.so`` behind the C ABI in ``include/heddle.h``.  This package is
the Python host side: ctypes bindings (``_lib``), an owner of a device
context (``engine``), input builders (``systems``), units and the set-up
RNG.
"""
from . import units  # noqa: F401
from ._lib import HeddleError, load  # noqa: F401

__version__ = "0.1.0"
