"""isce3_amd: MI355X-native time-domain SAR backprojection.

A drop-in for ``isce3.focus.backproject`` whose work runs in hand-written HIP
kernels for gfx950 behind the C-ABI of include/isce3_amd_backproject.h.
The look-alike host types mirror isce3's module layout:

    isce3_amd.core       DateTime, StateVector, Orbit, LUT2d, KnabKernel, ...
    isce3_amd.geometry   DEMInterpolator
    isce3_amd.product    RadarGridParameters
    isce3_amd.container  RadarGeometry
    isce3_amd.focus      backproject, backproject_blocks, Scene, RangeComp,
                         form_linear_chirp
"""
from . import container, core, errors, focus, geometry, product, rangecomp  # noqa: F401

__version__ = "0.1.0"
