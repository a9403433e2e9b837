"""Adapters from isce3's own Python objects to this package's look-alikes.

A workflow that already holds isce3 objects (isce3.core.Orbit, LUT2d,
isce3.product.RadarGridParameters, isce3.container.RadarGeometry,
isce3.geometry.DEMInterpolator, isce3.core.TabulatedKernelF32) converts them
here and calls isce3_amd.focus.backproject.  Only public attributes of the
isce3 bindings are read (pybind core/Orbit.cpp:52-92, core/LUT2d.cpp,
product/RadarGridParameters.cpp:60-110, geometry/DEMInterpolator.cpp:24-110),
so isce3 itself is not imported.
"""
from __future__ import annotations

import numpy as np

from . import core, errors
from .container import RadarGeometry
from .geometry import DEMInterpolator
from .product import RadarGridParameters


def datetime_from(dt) -> core.DateTime:
    if isinstance(dt, core.DateTime):
        return dt
    return core.DateTime(dt.isoformat())


def orbit_from(orbit) -> core.Orbit:
    t = orbit.time
    method = getattr(orbit, "interp_method", core.OrbitInterpMethod.HERMITE)
    name = getattr(method, "name", str(method)).upper()
    m = core.OrbitInterpMethod.LEGENDRE if "LEGENDRE" in name else core.OrbitInterpMethod.HERMITE
    return core.Orbit.from_arrays(datetime_from(orbit.reference_epoch), float(t.first),
                                  float(t.spacing), np.asarray(orbit.position),
                                  np.asarray(orbit.velocity), m)


def lut2d_from(lut) -> core.LUT2d:
    if not getattr(lut, "have_data", False):
        out = core.LUT2d()
        out.ref_value = float(getattr(lut, "ref_value", 0.0))
        return out
    return core.LUT2d(float(lut.x_start), float(lut.y_start), float(lut.x_spacing),
                      float(lut.y_spacing), np.asarray(lut.data, dtype=np.float64))


def grid_from(grid) -> RadarGridParameters:
    side = grid.lookside
    side = getattr(side, "name", side)
    return RadarGridParameters(grid.sensing_start, grid.wavelength, grid.prf,
                               grid.starting_range, grid.range_pixel_spacing, str(side),
                               grid.length, grid.width, datetime_from(grid.ref_epoch))


def geometry_from(geom) -> RadarGeometry:
    return RadarGeometry(grid_from(geom.radar_grid), orbit_from(geom.orbit),
                         lut2d_from(geom.doppler))


def dem_from(dem) -> DEMInterpolator:
    out = DEMInterpolator(dem.ref_height, method=getattr(dem, "interp_method", "bilinear"),
                          epsg=int(getattr(dem, "epsg_code", 4326)))
    if getattr(dem, "have_raster", False):
        out.load_dem_array(np.asarray(dem.data, dtype=np.float32), dem.x_start, dem.y_start,
                           dem.delta_x, dem.delta_y, int(dem.epsg_code))
    return out


def kernel_from(kernel, knab_width: float | None = None, knab_bandwidth: float | None = None,
                table_size: int = 2048) -> core.TabulatedKernelF32:
    """Rebuild a TabulatedKernelF32.  With the Knab parameters the table is
    recomputed exactly; otherwise the isce3 kernel is sampled at the table
    nodes through its __call__ (equal to the table up to float rounding of
    the reference's 1/dx)."""
    if knab_width is not None:
        return core.TabulatedKernelF32(core.KnabKernel(knab_width, knab_bandwidth), table_size)
    if not callable(kernel):
        raise errors.InvalidArgument("kernel must be callable or Knab parameters given")
    return core.TabulatedKernelF32(_Sampled(kernel), table_size)


class _Sampled(core.Kernel):
    def __init__(self, k):
        super().__init__(float(k.width))
        self._k = k

    def __call__(self, x):
        return float(self._k(x))
