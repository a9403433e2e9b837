"""ctypes binding of ``libisce3_amd.so`` (declared in include/isce3_amd_backproject.h).

The shared library is built in-tree by ``isce3_amd._build`` (hipcc, gfx950).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import errors

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libisce3_amd.so")

# status codes (isce3::error::ErrorCode values plus negative argument errors)
BP_SUCCESS = 0
BP_FAILED_TO_CONVERGE = 6
BP_ERR_INVALID_ARGUMENT = -1
BP_ERR_OUT_OF_RANGE = -2
BP_ERR_DOMAIN = -3
BP_ERR_HIP = -4
BP_ERR_RUNTIME = -5
BP_ERR_OUT_OF_MEMORY = -6
BP_ERR_LENGTH = -7

PRECISION_PLAIN = 0
PRECISION_SPLIT = 1


class Orbit_t(C.Structure):
    _fields_ = [("t0", C.c_double), ("dt", C.c_double), ("n", C.c_int32),
                ("method", C.c_int32), ("pos", C.c_void_p), ("vel", C.c_void_p)]


class Lut2d_t(C.Structure):
    _fields_ = [("have_data", C.c_int32), ("bounds_error", C.c_int32),
                ("ref_value", C.c_double), ("xstart", C.c_double), ("ystart", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("width", C.c_int64),
                ("length", C.c_int64), ("data", C.c_void_p)]


class Grid_t(C.Structure):
    _fields_ = [("sensing_start", C.c_double), ("prf", C.c_double),
                ("starting_range", C.c_double), ("range_pixel_spacing", C.c_double),
                ("length", C.c_int64), ("width", C.c_int64), ("look_side", C.c_int32),
                ("pad_", C.c_int32)]


class RadarGeometry_t(C.Structure):
    _fields_ = [("grid", Grid_t), ("orbit", Orbit_t), ("doppler", Lut2d_t)]


class Acorr_t(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("width", C.c_double)]


class StatsF32_t(C.Structure):
    _fields_ = [("min", C.c_float), ("max", C.c_float), ("mean", C.c_float),
                ("pad_", C.c_int32), ("n_valid", C.c_int64), ("real_valued_mean", C.c_double),
                ("square_diff_sum", C.c_double)]


class StatsRealImagF32_t(C.Structure):
    _fields_ = [("real", StatsF32_t), ("imag", StatsF32_t), ("n_valid", C.c_int64)]


class Dem_t(C.Structure):
    _fields_ = [("have_raster", C.c_int32), ("epsg", C.c_int32), ("method", C.c_int32),
                ("pad_", C.c_int32), ("ref_height", C.c_double), ("xstart", C.c_double),
                ("ystart", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("width", C.c_int64), ("length", C.c_int64), ("data", C.c_void_p)]


class Kernel_t(C.Structure):
    _fields_ = [("halfwidth", C.c_double), ("n", C.c_int32), ("pad_", C.c_int32),
                ("table", C.c_void_p)]


class Rdr2GeoParams_t(C.Structure):
    _fields_ = [("tol_height", C.c_double), ("look_min", C.c_double), ("look_max", C.c_double)]


class Geo2RdrParams_t(C.Structure):
    _fields_ = [("tol_aztime", C.c_double), ("have_time_start", C.c_int32),
                ("have_time_end", C.c_int32), ("time_start", C.c_double),
                ("time_end", C.c_double)]


class Options_t(C.Structure):
    _fields_ = [("tile_rows", C.c_int32), ("chunk", C.c_int32), ("fit_degree", C.c_int32),
                ("order", C.c_int32), ("blocks_per_cu", C.c_int32), ("staged", C.c_int32),
                ("lane_rows", C.c_int32), ("reserved", C.c_int32 * 1)]


class OutBlock_t(C.Structure):
    _fields_ = [("grid", Grid_t), ("out_dev", C.c_void_p), ("height_dev", C.c_void_p),
                ("failed", C.c_int32), ("pad_", C.c_int32)]


class Stats_t(C.Structure):
    _fields_ = [("updates", C.c_int64), ("geometry_ms", C.c_double), ("sum_ms", C.c_double),
                ("any_failed", C.c_int32), ("pad_", C.c_int32), ("step_ms", C.c_double)]


class HostStats_t(C.Structure):
    _fields_ = [("h2d_bytes", C.c_int64), ("window_pulses", C.c_int64), ("bands", C.c_int32),
                ("pad_", C.c_int32), ("copy_ms", C.c_double)]


_P = C.POINTER
_SIGNATURES = {
    "bp_context_create": (C.c_int, [C.c_int, _P(C.c_void_p)]),
    "bp_context_destroy": (C.c_int, [C.c_void_p]),
    "bp_last_error_message": (C.c_char_p, []),
    "bp_scratch_bytes_per_pixel": (C.c_int64, [C.c_int]),
    "bp_backproject": (C.c_int, [C.c_void_p, _P(RadarGeometry_t), _P(RadarGeometry_t), _P(Dem_t),
                                 C.c_double, C.c_double, _P(Kernel_t), C.c_int,
                                 _P(Rdr2GeoParams_t), _P(Geo2RdrParams_t), C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_int, _P(Options_t), C.c_void_p,
                                 _P(C.c_int), _P(Stats_t)]),
    "bp_backproject_noalloc": (C.c_int, [C.c_void_p, _P(RadarGeometry_t), _P(RadarGeometry_t),
                                         _P(Dem_t), C.c_double, C.c_double, _P(Kernel_t), C.c_int,
                                         _P(Rdr2GeoParams_t), _P(Geo2RdrParams_t), C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_int, _P(Options_t),
                                         C.c_void_p, _P(C.c_int), _P(Stats_t)]),
    "bp_backproject_host": (C.c_int, [C.c_void_p, _P(RadarGeometry_t), _P(RadarGeometry_t),
                                      _P(Dem_t), C.c_double, C.c_double, _P(Kernel_t), C.c_int,
                                      _P(Rdr2GeoParams_t), _P(Geo2RdrParams_t), C.c_void_p,
                                      C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_int,
                                      _P(Options_t), C.c_void_p, _P(C.c_int), _P(Stats_t),
                                      _P(HostStats_t)]),
    "bp_backproject_blocks": (C.c_int, [C.c_void_p, _P(Orbit_t), _P(Lut2d_t), _P(OutBlock_t),
                                        C.c_int, _P(RadarGeometry_t), _P(Dem_t), C.c_double,
                                        C.c_double, _P(Kernel_t), C.c_int, _P(Rdr2GeoParams_t),
                                        _P(Geo2RdrParams_t), C.c_void_p, C.c_int, _P(Options_t),
                                        C.c_void_p, _P(C.c_int), _P(Stats_t)]),
    "bp_scene_create": (C.c_int, [C.c_void_p, _P(Orbit_t), _P(Lut2d_t), _P(RadarGeometry_t),
                                  _P(Dem_t), C.c_double, C.c_double, _P(Kernel_t), C.c_int,
                                  _P(Rdr2GeoParams_t), _P(Geo2RdrParams_t), C.c_void_p, C.c_int,
                                  C.c_int64, C.c_int, _P(C.c_void_p)]),
    "bp_scene_destroy": (C.c_int, [C.c_void_p]),
    "bp_scene_focus": (C.c_int, [C.c_void_p, _P(OutBlock_t), C.c_int, _P(Options_t), C.c_void_p,
                                 _P(C.c_int), _P(Stats_t)]),
    "bp_server_create": (C.c_int, [C.c_void_p, _P(Options_t), C.c_int, _P(C.c_void_p)]),
    "bp_server_destroy": (C.c_int, [C.c_void_p]),
    "bp_serve_submit": (C.c_int, [C.c_void_p, _P(OutBlock_t), _P(C.c_uint64)]),
    "bp_serve_poll": (C.c_int, [C.c_void_p, C.c_uint64, _P(C.c_int), _P(C.c_int)]),
    "bp_serve_wait": (C.c_int, [C.c_void_p, C.c_uint64, _P(C.c_int)]),
    "bp_rangecomp_create": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                      _P(C.c_void_p)]),
    "bp_rangecomp_destroy": (C.c_int, [C.c_void_p]),
    "bp_rangecomp_sizes": (C.c_int, [C.c_void_p, _P(C.c_int32)]),
    "bp_rangecomp_plan": (C.c_int, [C.c_void_p, _P(C.c_int32)]),
    "bp_rangecomp_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "bp_rslc_finisher_create": (C.c_int, [C.c_int, C.c_int64, _P(C.c_void_p)]),
    "bp_rslc_finisher_destroy": (C.c_int, [C.c_void_p]),
    "bp_rslc_finish_block": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                       C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                       _P(StatsRealImagF32_t), C.c_void_p]),
    "bp_acorr_eval": (C.c_double, [_P(Acorr_t), C.c_double]),
    "bp_presum_weights": (C.c_int, [_P(Acorr_t), C.c_void_p, C.c_int64, C.c_double,
                                    _P(C.c_int64), C.c_void_p, C.c_int64, _P(C.c_int64)]),
    "bp_presum_apply_weights": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]),
    "bp_presum_compute_ids": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]),
    "bp_presum_resample": (C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                     _P(Grid_t), _P(Orbit_t), _P(Lut2d_t), C.c_double,
                                     C.c_void_p, C.c_void_p, C.c_void_p]),
    "bp_orbit_interpolate": (C.c_int, [_P(Orbit_t), C.c_double, C.c_int, _P(C.c_double),
                                       _P(C.c_double)]),
    "bp_lonlat_to_xyz": (None, [C.c_double, C.c_double, _P(C.c_double), _P(C.c_double)]),
    "bp_xyz_to_lonlat": (None, [C.c_double, C.c_double, _P(C.c_double), _P(C.c_double)]),
    "bp_dem_interpolate_lonlat": (C.c_double, [_P(Dem_t), C.c_double, C.c_double]),
    "bp_dem_interpolate_xy": (C.c_double, [_P(Dem_t), C.c_double, C.c_double]),
    "bp_proj_forward": (C.c_int, [C.c_int, _P(C.c_double), _P(C.c_double)]),
    "bp_proj_inverse": (C.c_int, [C.c_int, _P(C.c_double), _P(C.c_double)]),
    "bp_lut2d_eval": (C.c_double, [_P(Lut2d_t), C.c_double, C.c_double]),
    "bp_sincos_turns": (None, [C.c_double, _P(C.c_double), _P(C.c_double)]),
    "bp_knab_eval": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "bp_tabulate_knab": (C.c_int, [C.c_double, C.c_double, C.c_int, _P(C.c_float)]),
    "bp_rdr2geo_bracket": (C.c_int, [C.c_double, C.c_double, C.c_double, _P(Orbit_t), _P(Dem_t),
                                     C.c_double, C.c_int, _P(Rdr2GeoParams_t), _P(C.c_double)]),
    "bp_geo2rdr_bracket": (C.c_int, [_P(C.c_double), _P(Orbit_t), _P(Lut2d_t), C.c_double,
                                     C.c_int, _P(Geo2RdrParams_t), _P(C.c_double),
                                     _P(C.c_double)]),
}

_lib = None


def lib():
    """Load the native library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"isce3_amd native library not found at {LIB_PATH}; build it with "
            "`python -m isce3_amd._build` (hipcc, gfx950)")
    handle = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().bp_last_error_message()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> int:
    """Raise the exception the reference raises for a negative status code."""
    if rc >= 0:
        return rc
    msg = last_error() or what
    if rc == BP_ERR_INVALID_ARGUMENT:
        raise errors.InvalidArgument(msg)
    if rc == BP_ERR_OUT_OF_RANGE:
        raise errors.OutOfRange(msg)
    if rc == BP_ERR_DOMAIN:
        raise errors.DomainError(msg)
    if rc == BP_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    if rc == BP_ERR_LENGTH:
        raise errors.LengthError(msg)
    if rc == BP_ERR_HIP:
        raise errors.HipError(msg)
    raise errors.RuntimeError_(msg)


def dptr(a: np.ndarray):
    return C.cast(a.ctypes.data, _P(C.c_double))
