"""MI355X-native FDTD hot path: Yee curl + CPML + dispersive ADE + TFSF + running DFT.

The compute path is HIP (fdtd_amd/csrc, built into libfdtd_hip.so); this
package is its Python binding.
"""
from .grid import Batch, Grid, FdtdError, load_library  # noqa: F401

__all__ = ["Batch", "Grid", "FdtdError", "load_library"]
