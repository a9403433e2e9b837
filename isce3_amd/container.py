"""RadarGeometry look-alike (isce3.container.RadarGeometry,
cxx/isce3/container/RadarGeometry.{h,icc})."""
from __future__ import annotations

from .core import LUT2d, Linspace, Orbit
from .product import RadarGridParameters


class RadarGeometry:
    """Radar grid + orbit + Doppler.  Like the reference constructor
    (container/RadarGeometry.icc:7-26), the grid's sensing start is re-expressed
    relative to the orbit's reference epoch when the two epochs differ."""

    def __init__(self, radar_grid: RadarGridParameters, orbit: Orbit, doppler: LUT2d):
        grid = radar_grid.copy()
        if grid.ref_epoch != orbit.reference_epoch:
            dt = (grid.ref_epoch - orbit.reference_epoch).total_seconds()
            grid.sensing_start = grid.sensing_start + dt
            grid.ref_epoch = orbit.reference_epoch
        self._grid = grid
        self._orbit = orbit
        self._doppler = doppler

    @property
    def radar_grid(self) -> RadarGridParameters:
        return self._grid

    @property
    def orbit(self) -> Orbit:
        return self._orbit

    @property
    def doppler(self) -> LUT2d:
        return self._doppler

    @property
    def reference_epoch(self):
        return self._orbit.reference_epoch

    @property
    def grid_length(self) -> int:
        return self._grid.length

    @property
    def grid_width(self) -> int:
        return self._grid.width

    @property
    def sensing_time(self) -> Linspace:
        return Linspace(self._grid.sensing_start, 1.0 / self._grid.prf, self._grid.length)

    @property
    def slant_range(self) -> Linspace:
        return Linspace(self._grid.starting_range, self._grid.range_pixel_spacing,
                        self._grid.width)

    @property
    def look_side(self):
        return self._grid.lookside
