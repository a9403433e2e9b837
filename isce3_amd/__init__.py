"""isce3_amd: MI355X-native time-domain SAR backprojection.

A drop-in for ``isce3.focus.backproject`` whose work runs in hand-written HIP
kernels for gfx950 behind the C-ABI of include/isce3_amd_backproject.h.
The look-alike host types mirror isce3's module layout:

    isce3_amd.core       DateTime, StateVector, Orbit, LUT2d, KnabKernel, ...
    isce3_amd.geometry   DEMInterpolator
    isce3_amd.product    RadarGridParameters
    isce3_amd.container  RadarGeometry
    isce3_amd.focus      backproject, backproject_blocks, Scene, RangeComp,
                         form_linear_chirp, get_sar_duration, get_radar_velocities
    isce3_amd.rslc       BackgroundWriter, StatsRealImagFloat32, to_complex32,
                         truncate_mantissa (block finishing on the GPU)
    isce3_amd.planning   plan_processing_blocks and the workflow's parameter helpers
    isce3_amd.presum     BLU presumming of dithered-PRF pulses (resample, weights)
"""
from . import (container, core, errors, focus, geometry, planning, presum,  # noqa: F401
               product, rangecomp, rslc)

__version__ = "0.1.0"
