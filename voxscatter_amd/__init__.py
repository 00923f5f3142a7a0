"""voxscatter_amd: the solve loop of VoxScatter (FFT-accelerated DDA volume-integral matvec, Chan
circulant preconditioners, BiCGStab/GMRES) on AMD Instinct MI355X (gfx950), behind a C ABI
(include/voxscatter.h) with a Python mirror of the reference's MATLAB functions (``api``).
"""
from . import geometry, excitation  # noqa: F401  (host-side set-up, no device work)

__all__ = ["geometry", "excitation", "api", "capi", "dda"]
