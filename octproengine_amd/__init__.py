"""octproengine_amd: an MI355X-native (gfx950) HIP backend for OCTproEngine.

Layers
  include/ope_hip.h          C ABI of libope_hip.so (the drop-in boundary)
  csrc/hip/                  fused chain kernels and the C ABI
  csrc/core, csrc/backends   C++ ope::Processor API and the HIP backend
  csrc/python                pybind11 module ``octproengine`` (same API as the
                             reference's Python bindings)
  _capi.py                   ctypes binding of the C ABI
  synthetic.py               seeded fringe generator

Native libraries are built in-tree by ``python -m octproengine_amd.build``.
"""

from pathlib import Path

__version__ = "0.1.0"
PACKAGE_DIR = Path(__file__).resolve().parent


def capi():
    """Returns the ctypes binding of libope_hip.so (raises if not built)."""
    from . import _capi
    return _capi
