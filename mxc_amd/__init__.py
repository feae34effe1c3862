"""MI355X-native H^2 solve path (H^2 operator, ULV preconditioner, GMRES).

The numeric path is HIP code for gfx950 behind the C ABI in include/mxc_gpu.h;
the host layer (tree, block lists, far-body sampling, H2MatrixSolver) is C++
behind include/mxc_h2.h. This package only loads and drives them.
"""
from . import _native  # noqa: F401

__all__ = ["_native"]
