"""RadarGridParameters look-alike (isce3.product.RadarGridParameters,
cxx/isce3/product/RadarGridParameters.h; pybind product/RadarGridParameters.cpp:32-80)."""
from __future__ import annotations

from .core import DateTime, LookSide, speed_of_light


class RadarGridParameters:
    """RadarGridParameters(sensing_start, wavelength, prf, starting_range,
    range_pixel_spacing, lookside, length, width, ref_epoch)."""

    def __init__(self, sensing_start: float, wavelength: float, prf: float,
                 starting_range: float, range_pixel_spacing: float, lookside, length: int,
                 width: int, ref_epoch: DateTime):
        self.sensing_start = float(sensing_start)
        self.wavelength = float(wavelength)
        self.prf = float(prf)
        self.starting_range = float(starting_range)
        self.range_pixel_spacing = float(range_pixel_spacing)
        self.lookside = LookSide.parse(lookside)
        self.length = int(length)
        self.width = int(width)
        self.ref_epoch = DateTime(ref_epoch)

    @property
    def look_side(self) -> LookSide:
        return self.lookside

    @property
    def az_time_interval(self) -> float:
        return 1.0 / self.prf

    @property
    def size(self) -> int:
        return self.length * self.width

    @property
    def shape(self):
        return (self.length, self.width)

    def sensing_time(self, line: float) -> float:
        return self.sensing_start + line / self.prf

    @property
    def sensing_stop(self) -> float:
        return self.sensing_time(self.length - 1.0)

    @property
    def sensing_mid(self) -> float:
        return 0.5 * (self.sensing_start + self.sensing_stop)

    def slant_range(self, sample: float) -> float:
        return self.starting_range + sample * self.range_pixel_spacing

    @property
    def end_range(self) -> float:
        return self.slant_range(self.width - 1.0)

    @property
    def mid_range(self) -> float:
        return 0.5 * (self.starting_range + self.end_range)

    @property
    def range_sampling_rate(self) -> float:
        return speed_of_light / (2.0 * self.range_pixel_spacing)

    def offset_and_resize(self, yoff: float, xoff: float, ysize: int, xsize: int):
        """RadarGridParameters.h offsetAndResize: a sub-grid starting at (yoff, xoff)."""
        return RadarGridParameters(self.sensing_time(yoff), self.wavelength, self.prf,
                                   self.slant_range(xoff), self.range_pixel_spacing,
                                   self.lookside, ysize, xsize, self.ref_epoch)

    def __getitem__(self, key) -> "RadarGridParameters":
        """Sub-grid of a pair of slices (pybind product/RadarGridParameters.cpp:126-151):
        a row step scales the PRF down, a column step scales the spacing up."""
        if not isinstance(key, tuple) or len(key) != 2:
            raise ValueError("require 2 slices")
        rows, cols = key
        if not isinstance(rows, slice) or not isinstance(cols, slice):
            raise TypeError("require 2 slices")
        start, stop, step = rows.indices(self.length)
        if step <= 0:
            raise IndexError("cannot reverse grid")
        nt = len(range(start, stop, step))
        prf = self.prf / step
        t0 = self.sensing_start + start / self.prf
        start, stop, step = cols.indices(self.width)
        if step <= 0:
            raise IndexError("cannot reverse grid")
        nr = len(range(start, stop, step))
        dr = self.range_pixel_spacing * step
        r0 = self.starting_range + start * self.range_pixel_spacing
        return RadarGridParameters(t0, self.wavelength, prf, r0, dr, self.lookside, nt, nr,
                                   self.ref_epoch)

    def copy(self) -> "RadarGridParameters":
        return RadarGridParameters(self.sensing_start, self.wavelength, self.prf,
                                   self.starting_range, self.range_pixel_spacing,
                                   self.lookside, self.length, self.width, self.ref_epoch)
