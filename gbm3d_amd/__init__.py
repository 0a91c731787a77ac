"""gbm3d_amd -- MI355X (gfx950) implementation of the GBM3D denoising hot path.

The compute runs in libgbm3d.so (hand-written HIP kernels, C ABI declared in
include/gbm3d.h).  This package is the host-side mirror of the reference's
MATLAB call surface plus the synthetic-workload generator.
"""
from .api import (GBM3D, GBM3D_distributed, Context, checkBM3Dopts, default_context, denoiseAndAggregateGPU,  # noqa: F401
                  denoiseAndAggregateWie, denoiseAndAggregateWieFast, getColoredTau, getColoredWie, getTau, matchBlocksCPU, matchBlocksGPU, matchBlocksWie, myPSNR, resolve_opts,
                  setBM3Dopts, validate)
from ._abi import GBM3DError, LIB_PATH, load_library  # noqa: F401
