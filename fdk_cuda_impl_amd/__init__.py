"""MI355X (gfx950) implementation of the YKCBCT flat-panel circular-orbit path:
FDK filter + voxel-driven backprojection, Joseph / Siddon forward projection,
matched transposes, the SIRT / SART / OS-SART and CGLS solvers, smoothed TV,
calibrated-geometry diagnostics and raw stage dumps.

The compute path is ``libyk_hip.so`` (hand-written HIP kernels behind the C ABI
in ``include/yk_hip.h``); this package is its host-side binding.  Importing it
does not load the library; the first call does, and fails loudly if the library
has not been built (``fdk_cuda_impl_amd.build.build()``).
"""
from .core import (Context, FdkOptions, FdkPlan, Projector, ScanGeometry, build_views, derive,
                   fdk_coefficients, filter_weights, geometry_diagnostics, grid_of, make_filter,
                   padded_n)
from ._capi import YkError

__all__ = ["Context", "FdkOptions", "FdkPlan", "Projector", "ScanGeometry", "build_views",
           "derive", "fdk_coefficients", "filter_weights", "geometry_diagnostics", "grid_of",
           "make_filter", "padded_n",
           "YkError"]
