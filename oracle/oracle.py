"""ctypes front end of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker of the HIP product.  It restates the reference CPU
backprojector (isce3 cxx/isce3/focus/Backproject.cpp:30-211); see
bp_oracle.cpp for the per-function citations.  It takes the same host-side
geometry objects as isce3_amd.focus.backproject (duck-typed: grids, orbits,
LUTs, DEMs and tabulated kernels expose plain attributes) and packs them into
its own C structs.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libisce3_ref_blocks.so")

MODE_REFERENCE = 0
MODE_FP64 = 1
SUCCESS = 0
FAILED_TO_CONVERGE = 6
ARG_ERROR = -1
OUT_OF_RANGE = -2


class Orbit_t(C.Structure):
    _fields_ = [("t0", C.c_double), ("dt", C.c_double), ("n", C.c_int32), ("method", C.c_int32),
                ("pos", C.c_void_p), ("vel", C.c_void_p)]


class Lut_t(C.Structure):
    _fields_ = [("have_data", C.c_int32), ("bounds_error", C.c_int32), ("ref_value", C.c_double),
                ("xstart", C.c_double), ("ystart", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("width", C.c_int64), ("length", C.c_int64),
                ("data", C.c_void_p)]


class Grid_t(C.Structure):
    _fields_ = [("sensing_start", C.c_double), ("prf", C.c_double),
                ("starting_range", C.c_double), ("range_pixel_spacing", C.c_double),
                ("length", C.c_int64), ("width", C.c_int64), ("look_side", C.c_int32),
                ("pad_", C.c_int32)]


DEM_METHODS = {"sinc": 0, "bilinear": 1, "bicubic": 2, "nearest": 3, "biquintic": 4}


class Dem_t(C.Structure):
    _fields_ = [("have_raster", C.c_int32), ("epsg", C.c_int32), ("method", C.c_int32),
                ("pad_", C.c_int32), ("ref_height", C.c_double),
                ("xstart", C.c_double), ("ystart", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("width", C.c_int64), ("length", C.c_int64),
                ("data", C.c_void_p)]


class Kernel_t(C.Structure):
    _fields_ = [("halfwidth", C.c_double), ("n", C.c_int32), ("pad_", C.c_int32),
                ("table", C.c_void_p), ("table64", C.c_void_p)]


class R2G_t(C.Structure):
    _fields_ = [("tol_height", C.c_double), ("look_min", C.c_double), ("look_max", C.c_double)]


class G2R_t(C.Structure):
    _fields_ = [("tol_aztime", C.c_double), ("have_time_start", C.c_int32),
                ("have_time_end", C.c_int32), ("time_start", C.c_double),
                ("time_end", C.c_double)]


class PixGeom_t(C.Structure):
    _fields_ = [("x", C.c_double * 3), ("tau_atm", C.c_double), ("kstart", C.c_int32),
                ("kstop", C.c_int32), ("status", C.c_int32), ("pad_", C.c_int32)]


_lib = None
_ref = None
_D = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"oracle not built: {LIB} (run `make -C oracle`)")
        h = C.CDLL(LIB)
        h.orc_last_error.restype = C.c_char_p
        for name in ("orc_sinc", "orc_knab"):
            getattr(h, name).restype = C.c_double
        h.orc_sinc.argtypes = [C.c_double]
        h.orc_knab.argtypes = [C.c_double, C.c_double, C.c_double]
        h.orc_tabulate_knab_f32.argtypes = [C.c_double, C.c_double, C.c_int, C.c_void_p]
        h.orc_tabulate_knab_f64.argtypes = [C.c_double, C.c_double, C.c_int, C.c_void_p]
        h.orc_tab_eval_f32.restype = C.c_float
        h.orc_tab_eval_f32.argtypes = [C.POINTER(Kernel_t), C.c_double]
        h.orc_tab_eval_f64.restype = C.c_double
        h.orc_tab_eval_f64.argtypes = [C.POINTER(Kernel_t), C.c_double]
        h.orc_interp1d_c64.argtypes = [C.POINTER(Kernel_t), C.c_void_p, C.c_int64, C.c_double,
                                       C.c_void_p]
        h.orc_linspace_search.restype = C.c_int
        h.orc_linspace_search.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double]
        h.orc_orbit_interpolate.restype = C.c_int
        h.orc_orbit_interpolate.argtypes = [C.POINTER(Orbit_t), C.c_double, C.c_int, _D, _D]
        h.orc_lonlat_to_xyz.argtypes = [C.c_double, C.c_double, _D, _D]
        h.orc_xyz_to_lonlat.argtypes = [C.c_double, C.c_double, _D, _D]
        h.orc_dem_interpolate_lonlat.restype = C.c_double
        h.orc_dem_interpolate_lonlat.argtypes = [C.POINTER(Dem_t), C.c_double, C.c_double]
        h.orc_proj_forward.restype = C.c_int
        h.orc_proj_forward.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        h.orc_proj_inverse.restype = C.c_int
        h.orc_proj_inverse.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        h.orc_interp2d_f32.restype = C.c_float
        h.orc_interp2d_f32.argtypes = [C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_int64,
                                       C.c_int64]
        h.orc_interp2d_f64.restype = C.c_double
        h.orc_interp2d_f64.argtypes = [C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_int64,
                                       C.c_int64]
        h.orc_dem_interpolate_xy.restype = C.c_double
        h.orc_dem_interpolate_xy.argtypes = [C.POINTER(Dem_t), C.c_double, C.c_double]
        h.orc_lut2d_eval.restype = C.c_double
        h.orc_lut2d_eval.argtypes = [C.POINTER(Lut_t), C.c_double, C.c_double]
        h.orc_bistatic_delay.restype = C.c_double
        h.orc_bistatic_delay.argtypes = [_D, _D, _D]
        h.orc_dry_tropo_tsx.restype = C.c_double
        h.orc_dry_tropo_tsx.argtypes = [_D, _D]
        h.orc_brent_testfn.restype = C.c_int
        h.orc_brent_testfn.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _D]
        h.orc_rdr2geo_bracket.restype = C.c_int
        h.orc_rdr2geo_bracket.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(Orbit_t),
                                          C.POINTER(Dem_t), C.c_double, C.c_int,
                                          C.POINTER(R2G_t), _D]
        h.orc_geo2rdr_bracket.restype = C.c_int
        h.orc_geo2rdr_bracket.argtypes = [_D, C.POINTER(Orbit_t), C.POINTER(Lut_t), C.c_double,
                                          C.c_int, C.POINTER(G2R_t), _D, _D]
        h.orc_backproject.restype = C.c_int
        h.orc_backproject.argtypes = [C.POINTER(Grid_t), C.POINTER(Orbit_t), C.POINTER(Lut_t),
                                      C.POINTER(Grid_t), C.POINTER(Orbit_t), C.POINTER(Lut_t),
                                      C.POINTER(Dem_t), C.c_double, C.c_double,
                                      C.POINTER(Kernel_t), C.c_int, C.POINTER(R2G_t),
                                      C.POINTER(G2R_t), C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int]
        _lib = h
    return _lib


def ref_lib():
    """Reference building blocks compiled from the isce3 sources (None if absent)."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        h = C.CDLL(REF_LIB)
        h.ref_sinc.restype = C.c_double
        h.ref_sinc.argtypes = [C.c_double]
        h.ref_knab.restype = C.c_double
        h.ref_knab.argtypes = [C.c_double, C.c_double, C.c_double]
        h.ref_tab_table_f32.argtypes = [C.c_double, C.c_double, C.c_int, C.c_void_p]
        h.ref_tab_table_f64.argtypes = [C.c_double, C.c_double, C.c_int, C.c_void_p]
        h.ref_tab_eval_f32.argtypes = [C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_int64,
                                       C.c_void_p]
        h.ref_interp1d_c64.argtypes = [C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_int64,
                                       C.c_void_p, C.c_int64, C.c_void_p]
        h.ref_linspace_search.restype = C.c_int
        h.ref_linspace_search.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double]
        FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
        h.ref_find_zero_brent.restype = C.c_int
        h.ref_find_zero_brent.argtypes = [C.c_double, C.c_double, FN, C.c_void_p, C.c_double, _D]
        h.FN = FN
        _ref = h
    return _ref


# ------------------------------------------------------------------ packing
def orbit_struct(orbit) -> Orbit_t:
    pos = np.ascontiguousarray(orbit.position, dtype=np.float64)
    vel = np.ascontiguousarray(orbit.velocity, dtype=np.float64)
    s = Orbit_t(orbit.time.first, orbit.time.spacing, orbit.time.size, int(orbit.interp_method),
                pos.ctypes.data, vel.ctypes.data)
    s._keep = (pos, vel)
    return s


def lut_struct(lut) -> Lut_t:
    if not lut.have_data:
        return Lut_t(0, int(lut.bounds_error), lut.ref_value, 0., 0., 1., 1., 0, 0, None)
    data = np.ascontiguousarray(lut.data, dtype=np.float64)
    s = Lut_t(1, int(lut.bounds_error), lut.ref_value, lut.x_start, lut.y_start, lut.x_spacing,
              lut.y_spacing, data.shape[1], data.shape[0], data.ctypes.data)
    s._keep = data
    return s


def grid_struct(geometry) -> Grid_t:
    g = geometry.radar_grid
    return Grid_t(g.sensing_start, g.prf, g.starting_range, g.range_pixel_spacing, g.length,
                  g.width, int(g.lookside), 0)


def dem_struct(dem) -> Dem_t:
    method = DEM_METHODS[str(getattr(dem, "interp_method", "bilinear")).lower()]
    if not dem.have_raster:
        return Dem_t(0, dem.epsg_code, method, 0, dem.ref_height, 0., 0., 1., 1., 0, 0, None)
    data = np.ascontiguousarray(dem.data, dtype=np.float32)
    s = Dem_t(1, dem.epsg_code, method, 0, dem.ref_height, dem.x_start, dem.y_start,
              dem.delta_x, dem.delta_y, data.shape[1], data.shape[0], data.ctypes.data)
    s._keep = data
    return s


def kernel_struct(kernel, table64: Optional[np.ndarray] = None) -> Kernel_t:
    tab = np.ascontiguousarray(kernel.table, dtype=np.float32)
    t64 = None if table64 is None else np.ascontiguousarray(table64, dtype=np.float64)
    s = Kernel_t(kernel.width / 2.0, tab.size, 0, tab.ctypes.data,
                 None if t64 is None else t64.ctypes.data)
    s._keep = (tab, t64)
    return s


def knab_table_f64(width: float, bandwidth: float, n: int) -> np.ndarray:
    t = np.empty(n, dtype=np.float64)
    lib().orc_tabulate_knab_f64(width, bandwidth, n, t.ctypes.data)
    return t


class OracleError(RuntimeError):
    pass


def backproject(out_geometry, signal_data: np.ndarray, in_geometry, dem, fc: float, ds: float,
                kernel, dry_tropo_model: str = "tsx", rdr2geo_params: Optional[dict] = None,
                geo2rdr_params: Optional[dict] = None, mode: int = MODE_REFERENCE,
                subset: Optional[np.ndarray] = None, nthreads: int = 0,
                kernel_bandwidth: Optional[float] = None, want_height: bool = False,
                want_diag: bool = False):
    """Run the oracle.  Returns (out, height, diag, status) where out is complex64
    [L, W] (or [len(subset)]); mode MODE_FP64 needs kernel_bandwidth to rebuild
    the Knab table in double."""
    L, W = out_geometry.grid_length, out_geometry.grid_width
    n = L * W if subset is None else len(subset)
    out = np.empty(n, dtype=np.complex64)
    height = np.empty(n, dtype=np.float32) if want_height else None
    diag = (PixGeom_t * n)() if want_diag else None
    sig = np.ascontiguousarray(signal_data, dtype=np.complex64)
    t64 = None
    if mode == MODE_FP64:
        if kernel_bandwidth is None:
            raise ValueError("FP64 mode needs kernel_bandwidth")
        t64 = knab_table_f64(kernel.width, kernel_bandwidth, kernel.table.size)
    r2g = R2G_t(1e-5, 0.0, np.pi / 2)
    for k, v in (rdr2geo_params or {}).items():
        setattr(r2g, k, float(v))
    g2r = G2R_t(1e-7, 0, 0, 0.0, 0.0)
    for k, v in (geo2rdr_params or {}).items():
        if k == "tol_aztime":
            g2r.tol_aztime = float(v)
        elif v is not None:
            setattr(g2r, "have_" + k, 1)
            setattr(g2r, k, float(v))
    tropo = {"nodelay": 0, "tsx": 1}[dry_tropo_model]
    og, ig = grid_struct(out_geometry), grid_struct(in_geometry)
    oo, io = orbit_struct(out_geometry.orbit), orbit_struct(in_geometry.orbit)
    od, idp = lut_struct(out_geometry.doppler), lut_struct(in_geometry.doppler)
    dm = dem_struct(dem)
    kr = kernel_struct(kernel, t64)
    sub = None if subset is None else np.ascontiguousarray(subset, dtype=np.int64)
    rc = lib().orc_backproject(C.byref(og), C.byref(oo), C.byref(od), C.byref(ig), C.byref(io),
                               C.byref(idp), C.byref(dm), float(fc), float(ds), C.byref(kr),
                               tropo, C.byref(r2g), C.byref(g2r), sig.ctypes.data,
                               out.ctypes.data,
                               None if height is None else height.ctypes.data, int(mode),
                               None if sub is None else sub.ctypes.data,
                               0 if sub is None else sub.size,
                               None if diag is None else C.addressof(diag), int(nthreads))
    if rc < 0:
        raise OracleError(lib().orc_last_error().decode())
    if subset is None:
        out = out.reshape(L, W)
        if height is not None:
            height = height.reshape(L, W)
    return out, height, diag, rc
