"""DEMInterpolator look-alike (isce3.geometry.DEMInterpolator).

The reference loads rasters through GDAL (geometry/DEMInterpolator.cpp:442-494),
which is outside this package; rasters are handed over as in-memory arrays
with :meth:`DEMInterpolator.load_dem_array` instead.  Lookup semantics follow
geometry/DEMInterpolator.cpp:593-662: (lon, lat) through the raster's map
projection (any EPSG code of core/Projections.cpp's factory), interpolation by
the nearest / bilinear / bicubic / biquintic method, reference height outside
the raster's [2, n-1) interior.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native, errors

# isce3::core::dataInterpMethod values (core/Constants.h:23-29)
_METHODS = {"sinc": 0, "bilinear": 1, "bicubic": 2, "nearest": 3, "biquintic": 4}


class DEMInterpolator:
    """DEMInterpolator(height=0., method="bilinear", epsg=4326)."""

    def __init__(self, height: float = 0.0, method: str = "bilinear", epsg: int = 4326):
        # the reference stores the reference height as float (DEMInterpolator.h:195)
        self._ref_height = float(np.float32(height))
        self._method = self._parse_method(method)
        self._epsg = int(epsg)
        self._data = None
        self._x_start = self._y_start = 0.0
        self._dx = self._dy = 1.0
        self._stats = (self._ref_height, self._ref_height, self._ref_height)

    @staticmethod
    def _parse_method(method) -> str:
        # core/Constants.cpp:10-24 (parseDataInterpMethod)
        m = str(getattr(method, "name", method)).lower()
        if m not in _METHODS:
            raise errors.InvalidArgument("Unknown interp method")
        return m

    def load_dem_array(self, data, x_start: float, y_start: float, delta_x: float,
                       delta_y: float, epsg: int = 4326):
        """Use an in-memory float raster: data[row, col] at (x_start + col*delta_x,
        y_start + row*delta_y) in EPSG `epsg` coordinates (degrees for 4326)."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        if data.ndim != 2:
            raise errors.InvalidArgument("DEM raster must be 2-D")
        from .core import make_projection
        make_projection(int(epsg))       # RuntimeError for a code the factory rejects
        self._data = data
        self._data.setflags(write=False)
        self._x_start, self._y_start = float(x_start), float(y_start)
        self._dx, self._dy = float(delta_x), float(delta_y)
        self._epsg = int(epsg)
        return self

    @property
    def ref_height(self) -> float:
        return self._ref_height

    @ref_height.setter
    def ref_height(self, h: float):
        self._ref_height = float(np.float32(h))

    @property
    def have_raster(self) -> bool:
        return self._data is not None

    @property
    def interp_method(self) -> str:
        return self._method

    @property
    def data(self) -> np.ndarray:
        if self._data is None:
            raise IndexError("Tried to access DEM data but size=0")
        return self._data

    @property
    def x_start(self) -> float:
        return self._x_start

    @property
    def y_start(self) -> float:
        return self._y_start

    @property
    def delta_x(self) -> float:
        return self._dx

    @property
    def delta_y(self) -> float:
        return self._dy

    @property
    def width(self) -> int:
        return 0 if self._data is None else int(self._data.shape[1])

    @property
    def length(self) -> int:
        return 0 if self._data is None else int(self._data.shape[0])

    @property
    def epsg_code(self) -> int:
        return self._epsg

    def compute_min_max_mean_height(self):
        if self._data is None:
            self._stats = (self._ref_height,) * 3
            return
        valid = self._data[np.isfinite(self._data)]
        if valid.size:
            self._stats = (float(valid.min()), float(valid.max()), float(valid.mean()))

    @property
    def min_height(self) -> float:
        return self._stats[0]

    @property
    def max_height(self) -> float:
        return self._stats[1]

    @property
    def mean_height(self) -> float:
        return self._stats[2]

    def _struct(self) -> _native.Dem_t:
        have = self._data is not None
        return _native.Dem_t(int(have), self._epsg, _METHODS[self._method], 0, self._ref_height,
                             self._x_start,
                             self._y_start, self._dx, self._dy, self.width, self.length,
                             self._data.ctypes.data if have else None)

    def interpolate_lonlat(self, lon: float, lat: float) -> float:
        s = self._struct()
        return _native.lib().bp_dem_interpolate_lonlat(C.byref(s), float(lon), float(lat))

    def interpolate_xy(self, x: float, y: float) -> float:
        s = self._struct()
        return _native.lib().bp_dem_interpolate_xy(C.byref(s), float(x), float(y))


def rdr2geo_bracket(aztime: float, slant_range: float, orbit, side, doppler: float,
                    wavelength: float, dem: "DEMInterpolator | None" = None,
                    tol_height: float = 1e-5, look_min: float = 0.0,
                    look_max: float = np.pi / 2) -> np.ndarray:
    """Radar (aztime, slant_range, Doppler) -> ECEF target on the DEM by Brent on
    the pseudo-look angle (pybind geometry/rdr2geo_roots.cpp:18-41,
    geometry/detail/Rdr2Geo.icc:172-242).  Raises RuntimeError on failure."""
    from .core import LookSide
    dem = dem if dem is not None else DEMInterpolator()
    side = LookSide.parse(side)
    o = orbit._struct()
    d = dem._struct()
    p = _native.Rdr2GeoParams_t(float(tol_height), float(look_min), float(look_max))
    xyz = np.empty(3)
    rc = _native.lib().bp_rdr2geo_bracket(float(aztime), float(slant_range), float(doppler),
                                          C.byref(o), C.byref(d), float(wavelength), int(side),
                                          C.byref(p), _native.dptr(xyz))
    if rc == 2:
        raise errors.OutOfRange("interpolation point outside orbit domain")
    if rc != 0:
        raise RuntimeError("failed to converge")
    return xyz


def geo2rdr_bracket(xyz, orbit, doppler, wavelength: float, side, tol_aztime: float = 1e-7,
                    time_start=None, time_end=None):
    """ECEF target -> (aztime, slant_range) by Brent on azimuth time against the
    Doppler LUT (pybind geometry/geo2rdr_roots.cpp:21-40,
    geometry/detail/Geo2Rdr.icc:182-238).  Raises RuntimeError on failure."""
    from .core import LookSide
    side = LookSide.parse(side)
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(3)
    o = orbit._struct()
    l = doppler._struct()
    p = _native.Geo2RdrParams_t(float(tol_aztime), int(time_start is not None),
                                int(time_end is not None),
                                float(time_start) if time_start is not None else 0.0,
                                float(time_end) if time_end is not None else 0.0)
    t = C.c_double()
    r = C.c_double()
    rc = _native.lib().bp_geo2rdr_bracket(_native.dptr(x), C.byref(o), C.byref(l),
                                          float(wavelength), int(side), C.byref(p),
                                          C.byref(t), C.byref(r))
    if rc != 0:
        raise RuntimeError("failed to converge")
    return t.value, r.value


def heading(lon: float, lat: float, vel) -> float:
    """Heading of velocity `vel` (ECEF) at (lon, lat), clockwise from north
    (geometry/geometry.cpp:399-408)."""
    import math
    v = np.asarray(vel, dtype=np.float64)
    coslat, sinlat = math.cos(lat), math.sin(lat)
    coslon, sinlon = math.cos(lon), math.sin(lon)
    return math.atan2(-sinlon * v[0] + coslon * v[1],
                      -sinlat * coslon * v[0] - sinlat * sinlon * v[1] + coslat * v[2])


def rdr2rdr(t: float, r: float, orbit, side, doppler, wavelength: float,
            dem: "DEMInterpolator | None" = None, ellipsoid=None, doppler_out=None,
            orbit_out=None, rdr2geo_params=None, geo2rdr_params=None):
    """Radar coordinates in one geometry -> another with a different Doppler
    or orbit: rdr2geo_bracket at the input Doppler, then geo2rdr_bracket
    against the output orbit and Doppler
    (python/packages/isce3/geometry/rdr2rdr.py:6-72)."""
    if orbit_out is None and doppler_out is None:
        return t, r
    orbit_out = orbit if orbit_out is None else orbit_out
    doppler_out = doppler if doppler_out is None else doppler_out
    dem = dem if dem is not None else DEMInterpolator()
    doppler_in = doppler.eval(t, r)
    xyz = rdr2geo_bracket(t, r, orbit, side, doppler_in, wavelength, dem,
                          **(rdr2geo_params or {}))
    return geo2rdr_bracket(xyz, orbit_out, doppler_out, wavelength, side,
                           **(geo2rdr_params or {}))
