"""Host-side look-alikes of the isce3.core types that cross the backprojection
boundary: DateTime/TimeDelta, StateVector, Orbit, LUT2d, Linspace, LookSide,
KnabKernel, TabulatedKernelF32 and Ellipsoid.

Constructor arguments, property names and error behaviour follow the isce3
pybind11 bindings (python/extensions/pybind_isce3/core/*.cpp).  The numerics
(orbit interpolation, Knab evaluation, table fill, LUT and ellipsoid
conversions) run in the native library, in the same FP64 code the GPU
kernels use.
"""
from __future__ import annotations

import ctypes as C
import datetime as _dt
import enum
import math
from typing import Sequence

import numpy as np

from . import _native, errors

speed_of_light = 299792458.0          # core/Constants.h:50
EarthSemiMajorAxis = 6378137.0        # core/Constants.h:41
EarthEccentricitySquared = 0.006694379990141317  # core/Constants.h:44


# --------------------------------------------------------------- time types
class TimeDelta:
    """Signed time interval in seconds (isce3::core::TimeDelta)."""

    __slots__ = ("_s",)

    def __init__(self, seconds: float = 0.0):
        self._s = float(seconds)

    def total_seconds(self) -> float:
        return self._s

    def __float__(self):
        return self._s

    def __add__(self, other):
        if isinstance(other, TimeDelta):
            return TimeDelta(self._s + other._s)
        return NotImplemented

    def __neg__(self):
        return TimeDelta(-self._s)

    def __eq__(self, other):
        return isinstance(other, TimeDelta) and self._s == other._s

    def __repr__(self):
        return f"TimeDelta({self._s!r})"


class DateTime:
    """Calendar instant with sub-second precision (isce3::core::DateTime).

    Stored as integer seconds since 1970-01-01T00:00:00 plus a fractional
    second in [0, 1), so differences of nearby instants keep full double
    precision.  Accepts an ISO-8601 string, a ``datetime.datetime`` or
    (year, month, day[, hour, minute, second[, frac]]).
    """

    __slots__ = ("_sec", "_frac")

    def __init__(self, *args):
        if len(args) == 1 and isinstance(args[0], DateTime):
            self._sec, self._frac = args[0]._sec, args[0]._frac
            return
        if len(args) == 1 and isinstance(args[0], str):
            s = args[0].strip()
            frac = 0.0
            if "." in s:
                s, f = s.split(".", 1)
                frac = float("0." + f)
            t = _dt.datetime.fromisoformat(s)
            self._set(t, frac)
            return
        if len(args) == 1 and isinstance(args[0], _dt.datetime):
            t = args[0]
            self._set(t.replace(microsecond=0), t.microsecond * 1e-6)
            return
        if 3 <= len(args) <= 7:
            y, mo, d = args[:3]
            h = args[3] if len(args) > 3 else 0
            mi = args[4] if len(args) > 4 else 0
            sec = args[5] if len(args) > 5 else 0
            frac = args[6] if len(args) > 6 else 0.0
            isec = int(math.floor(sec))
            frac += sec - isec
            self._set(_dt.datetime(int(y), int(mo), int(d), int(h), int(mi), isec), frac)
            return
        raise TypeError("invalid datetime object")

    def _set(self, t: _dt.datetime, frac: float):
        epoch = _dt.datetime(1970, 1, 1)
        sec = int((t - epoch).total_seconds())
        extra = math.floor(frac)
        self._sec = sec + int(extra)
        self._frac = float(frac - extra)

    @classmethod
    def _raw(cls, sec: int, frac: float) -> "DateTime":
        obj = cls.__new__(cls)
        extra = math.floor(frac)
        obj._sec = int(sec) + int(extra)
        obj._frac = float(frac - extra)
        return obj

    def __sub__(self, other):
        if isinstance(other, DateTime):
            return TimeDelta((self._sec - other._sec) + (self._frac - other._frac))
        if isinstance(other, TimeDelta):
            return self + (-other)
        return NotImplemented

    def __add__(self, other):
        if isinstance(other, (int, float)):
            other = TimeDelta(other)
        if isinstance(other, TimeDelta):
            s = other.total_seconds()
            isec = math.floor(s)
            return DateTime._raw(self._sec + int(isec), self._frac + (s - isec))
        return NotImplemented

    def __eq__(self, other):
        return (isinstance(other, DateTime) and self._sec == other._sec
                and self._frac == other._frac)

    def __ne__(self, other):
        return not self.__eq__(other)

    def __hash__(self):
        return hash((self._sec, self._frac))

    def is_close(self, other: "DateTime", tol: float = 1e-9) -> bool:
        return abs((self - other).total_seconds()) <= tol

    def isoformat(self) -> str:
        t = _dt.datetime(1970, 1, 1) + _dt.timedelta(seconds=self._sec)
        return t.strftime("%Y-%m-%dT%H:%M:%S") + f"{self._frac:.9f}"[1:]

    def __repr__(self):
        return f"DateTime('{self.isoformat()}')"


class StateVector:
    """Orbit sample: epoch, ECEF position (m) and velocity (m/s)."""

    def __init__(self, datetime: DateTime, position, velocity):
        self.datetime = DateTime(datetime)
        self.position = np.asarray(position, dtype=np.float64).reshape(3)
        self.velocity = np.asarray(velocity, dtype=np.float64).reshape(3)


class OrbitInterpMethod(enum.IntEnum):
    HERMITE = 0
    LEGENDRE = 1


class OrbitInterpBorderMode(enum.IntEnum):
    ERROR = 0
    EXTRAPOLATE = 1
    FILL_NAN = 2


class LookSide(enum.IntEnum):
    """isce3::core::LookSide values."""
    Left = 1
    Right = -1

    @staticmethod
    def parse(value) -> "LookSide":
        if isinstance(value, LookSide):
            return value
        if isinstance(value, str):
            v = value.strip().lower()
            if v == "left":
                return LookSide.Left
            if v == "right":
                return LookSide.Right
            raise errors.InvalidArgument(f"Invalid look side: {value}")
        return LookSide(int(value))


class Linspace:
    """Uniform sampling first + i*spacing, i in [0, size) (core/Linspace.h)."""

    def __init__(self, first: float, spacing: float, size: int):
        self.first = float(first)
        self.spacing = float(spacing)
        self.size = int(size)

    @property
    def last(self) -> float:
        return self[self.size - 1]

    def __getitem__(self, i: int) -> float:
        return self.first + i * self.spacing

    def __len__(self):
        return self.size

    def __array__(self, dtype=None, copy=None):
        a = self.first + np.arange(self.size, dtype=np.float64) * self.spacing
        return a.astype(dtype) if dtype is not None else a

    def search(self, val: float) -> int:
        """core/Linspace.icc:77-89."""
        if self.spacing >= 0:
            if val < self.first:
                return 0
            if val > self.last:
                return self.size
        else:
            if val > self.first:
                return 0
            if val < self.last:
                return self.size
        return int((val - self.first) / self.spacing + 1)

    def __eq__(self, other):
        return (isinstance(other, Linspace) and self.first == other.first
                and self.spacing == other.spacing and self.size == other.size)


# --------------------------------------------------------------------- orbit
class Orbit:
    """Uniformly sampled platform trajectory (isce3::core::Orbit).

    Orbit(state_vectors, interp_method=HERMITE, type="") uses the first state
    vector's epoch as reference epoch; Orbit(state_vectors, date_time,
    interp_method, type) sets it explicitly (pybind core/Orbit.cpp:41-51).
    """

    def __init__(self, state_vectors: Sequence[StateVector], *args, **kwargs):
        ref = kwargs.pop("date_time", None)
        args = list(args)
        if args and isinstance(args[0], DateTime):
            ref = args.pop(0)
        interp = kwargs.pop("interp_method", args.pop(0) if args else OrbitInterpMethod.HERMITE)
        self.type = kwargs.pop("type", args.pop(0) if args else "")
        if kwargs:
            raise TypeError(f"unexpected arguments {sorted(kwargs)}")
        svs = list(state_vectors)
        if len(svs) < 2:
            raise errors.InvalidArgument("at least two state vectors are required")
        if ref is None:
            ref = svs[0].datetime
        self._epoch = DateTime(ref)
        spacing = (svs[1].datetime - svs[0].datetime).total_seconds()
        for i in range(len(svs) - 1):
            d = (svs[i + 1].datetime - svs[i].datetime).total_seconds()
            if abs(d - spacing) > 1e-6:
                raise errors.InvalidArgument(
                    "non-uniform spacing between state vectors encountered - interval "
                    f"between state vector at position {i} and state vector at position "
                    f"{i + 1} is {d} s, expected {spacing} s")
        t0 = (svs[0].datetime - self._epoch).total_seconds()
        self._time = Linspace(t0, spacing, len(svs))
        self._pos = np.ascontiguousarray([sv.position for sv in svs], dtype=np.float64)
        self._vel = np.ascontiguousarray([sv.velocity for sv in svs], dtype=np.float64)
        self.interp_method = OrbitInterpMethod(int(interp))

    @classmethod
    def from_arrays(cls, reference_epoch: DateTime, t0: float, spacing: float, position,
                    velocity, interp_method=OrbitInterpMethod.HERMITE, type: str = "") -> "Orbit":
        """Build from uniformly spaced arrays (times relative to reference_epoch)."""
        pos = np.asarray(position, dtype=np.float64)
        vel = np.asarray(velocity, dtype=np.float64)
        svs = [StateVector(reference_epoch + TimeDelta(t0 + i * spacing), pos[i], vel[i])
               for i in range(pos.shape[0])]
        orb = cls(svs, reference_epoch, interp_method, type)
        orb._time = Linspace(t0, spacing, pos.shape[0])
        return orb

    @property
    def reference_epoch(self) -> DateTime:
        return self._epoch

    def update_reference_epoch(self, reference_epoch: DateTime):
        """core/Orbit.cpp:59-69: shift time tags so absolute times are unchanged."""
        dt = (self._epoch - reference_epoch).total_seconds()
        self._time = Linspace(self._time.first + dt, self._time.spacing, self._time.size)
        self._epoch = DateTime(reference_epoch)

    def copy(self) -> "Orbit":
        o = Orbit.__new__(Orbit)
        o._epoch = DateTime(self._epoch)
        o._time = Linspace(self._time.first, self._time.spacing, self._time.size)
        o._pos = self._pos.copy()
        o._vel = self._vel.copy()
        o.interp_method = self.interp_method
        o.type = self.type
        return o

    @property
    def time(self) -> Linspace:
        return self._time

    @property
    def position(self) -> np.ndarray:
        return self._pos

    @property
    def velocity(self) -> np.ndarray:
        return self._vel

    @property
    def spacing(self) -> float:
        return self._time.spacing

    @property
    def size(self) -> int:
        return self._time.size

    @property
    def start_time(self) -> float:
        return self._time.first

    @property
    def mid_time(self) -> float:
        return self.start_time + 0.5 * (self.size - 1) * self.spacing

    @property
    def end_time(self) -> float:
        return self._time.last

    @property
    def start_datetime(self) -> DateTime:
        return self._epoch + TimeDelta(self.start_time)

    @property
    def mid_datetime(self) -> DateTime:
        return self._epoch + TimeDelta(self.mid_time)

    @property
    def end_datetime(self) -> DateTime:
        return self._epoch + TimeDelta(self.end_time)

    def contains(self, time: float) -> bool:
        return self.start_time <= time <= self.end_time

    def _struct(self) -> _native.Orbit_t:
        return _native.Orbit_t(self._time.first, self._time.spacing, self._time.size,
                               int(self.interp_method), self._pos.ctypes.data,
                               self._vel.ctypes.data)

    def interpolate(self, t: float, border_mode=OrbitInterpBorderMode.ERROR):
        """Position and velocity at time t (core/Orbit.cpp:71-86)."""
        pos = np.empty(3)
        vel = np.empty(3)
        s = self._struct()
        rc = _native.lib().bp_orbit_interpolate(C.byref(s), float(t), int(border_mode),
                                                _native.dptr(pos), _native.dptr(vel))
        if rc != 0 and int(border_mode) == OrbitInterpBorderMode.ERROR:
            msg = {1: "insufficient orbit state vectors to form interpolant",
                   2: "interpolation point outside orbit domain",
                   3: "unexpected orbit interpolation method"}.get(rc, "orbit error")
            raise errors.OutOfRange(msg)
        return pos, vel


# --------------------------------------------------------------------- LUT2d
class LUT2d:
    """Bilinear 2-D look-up table of doubles (isce3::core::LUT2d<double>).

    LUT2d() is the constant-zero table (e.g. zero Doppler).  LUT2d(xstart,
    ystart, dx, dy, data, method="bilinear", b_error=True) or LUT2d(xcoord,
    ycoord, data, method, b_error) build a table (pybind core/LUT2d.cpp:36-90).
    x is slant range, y is azimuth time.
    """

    def __init__(self, *args, method="bilinear", b_error=True):
        self.have_data = False
        self.ref_value = 0.0
        self.x_start = self.y_start = 0.0
        self.x_spacing = self.y_spacing = 1.0
        self.data = np.zeros((0, 0))
        self.bounds_error = bool(b_error)
        if str(method).lower() != "bilinear":
            raise errors.InvalidArgument(f"unsupported LUT2d interpolation method {method!r}")
        self.interp_method = "bilinear"
        if not args:
            return
        if len(args) == 5:
            xstart, ystart, dx, dy, data = args
            data = np.ascontiguousarray(data, dtype=np.float64)
            if data.ndim != 2:
                raise errors.RuntimeError_("buffer object must be 2-D")
            self._set(float(xstart), float(ystart), float(dx), float(dy), data)
        elif len(args) == 3:
            xc, yc, data = (np.asarray(a, dtype=np.float64) for a in args)
            data = np.ascontiguousarray(data)
            if xc.size != data.shape[1] or yc.size != data.shape[0]:
                raise errors.InvalidArgument("Inconsistent shapes between data and coordinates")
            if xc.size == 1 and yc.size == 1:
                self.ref_value = float(data[0, 0])
                return
            if xc.size < 2 or yc.size < 2:
                raise errors.InvalidArgument("invalid LUT2d data shape")
            self._set(float(xc[0]), float(yc[0]), float(xc[1] - xc[0]), float(yc[1] - yc[0]), data)
        else:
            raise TypeError("LUT2d(), LUT2d(xstart, ystart, dx, dy, data) or LUT2d(x, y, data)")

    def _set(self, xstart, ystart, dx, dy, data):
        self.x_start, self.y_start, self.x_spacing, self.y_spacing = xstart, ystart, dx, dy
        self.data = data
        self.have_data = True
        self.ref_value = float(data[0, 0])

    @property
    def width(self) -> int:
        return int(self.data.shape[1]) if self.have_data else 0

    @property
    def length(self) -> int:
        return int(self.data.shape[0]) if self.have_data else 0

    @property
    def x_end(self) -> float:
        return self.x_start + self.x_spacing * (self.width - 1.0)

    @property
    def y_end(self) -> float:
        return self.y_start + self.y_spacing * (self.length - 1.0)

    def contains(self, y: float, x: float) -> bool:
        if not self.have_data:
            return True
        i = (x - self.x_start) / self.x_spacing
        j = (y - self.y_start) / self.y_spacing
        return 0.0 <= i <= self.width - 1.0 and 0.0 <= j <= self.length - 1.0

    def _struct(self) -> _native.Lut2d_t:
        return _native.Lut2d_t(int(self.have_data), int(self.bounds_error), self.ref_value,
                               self.x_start, self.y_start, self.x_spacing, self.y_spacing,
                               self.width, self.length,
                               self.data.ctypes.data if self.have_data else None)

    def eval(self, y: float, x: float) -> float:
        s = self._struct()
        return _native.lib().bp_lut2d_eval(C.byref(s), float(y), float(x))


# -------------------------------------------------------------------- kernels
class Kernel:
    """A function on [-width/2, width/2] (isce3::core::Kernel)."""

    def __init__(self, width: float):
        self._halfwidth = abs(float(width) / 2.0)

    @property
    def width(self) -> float:
        return self._halfwidth * 2


class KnabKernel(Kernel):
    """Knab (1983) interpolation kernel, double precision (core/Kernels.h:64-84)."""

    def __init__(self, width: float, bandwidth: float):
        super().__init__(width)
        self._bandwidth = float(bandwidth)

    @property
    def bandwidth(self) -> float:
        return self._bandwidth

    def __call__(self, x: float) -> float:
        if not (0.0 < self._bandwidth < 1.0):
            raise errors.RuntimeError_("Require 0 < bandwidth < 1")
        return _native.lib().bp_knab_eval(self.width, self._bandwidth, float(x))


class BartlettKernel(Kernel):
    """Triangle of the given width (core/Kernels.h:43-57, Kernels.icc:16-23)."""

    _kind = 1

    def _struct(self) -> _native.Acorr_t:
        return _native.Acorr_t(self._kind, 0, self.width)

    def __call__(self, x: float) -> float:
        return _native.lib().bp_acorr_eval(C.byref(self._struct()), float(x))


class AzimuthKernel(BartlettKernel):
    """Autocorrelation of the sinc^2 azimuth antenna pattern, a piecewise
    cubic of width 2 scale (core/Kernels.h:114-129, Kernels.icc:95-107)."""

    _kind = 0

    def __init__(self, scale: float):
        super().__init__(2.0 * float(scale))


class KernelF32(Kernel):
    """Base of single-precision kernels accepted by backproject."""


class TabulatedKernelF32(KernelF32):
    """Float look-up table of an even kernel (core/Kernels.h:131-151).

    TabulatedKernelF32(kernel, table_size): the table holds kernel(x) at
    table_size samples of x in [0, width/2]; evaluation interpolates
    linearly with the reference's index arithmetic.
    """

    def __init__(self, kernel: Kernel, table_size: int):
        super().__init__(kernel.width)
        n = int(table_size)
        if n < 2:
            raise errors.LengthError("Require table size >= 2.")
        self._table = np.empty(n, dtype=np.float32)
        if isinstance(kernel, KnabKernel):
            rc = _native.lib().bp_tabulate_knab(
                kernel.width, kernel.bandwidth, n,
                self._table.ctypes.data_as(C.POINTER(C.c_float)))
            _native.check(rc)
        elif isinstance(kernel, TabulatedKernelF32):
            hw = self._halfwidth
            dx = hw / (n - 1.0)
            self._table[:] = [kernel(i * dx) for i in range(n)]
        elif callable(kernel):
            hw = self._halfwidth
            dx = hw / (n - 1.0)
            self._table[:] = np.array([kernel(i * dx) for i in range(n)], dtype=np.float64)
        else:
            raise TypeError("kernel must be callable")
        self._table.setflags(write=False)

    @property
    def table(self) -> np.ndarray:
        return self._table

    def __call__(self, x: float) -> float:
        """core/Kernels.icc:137-154 (float table, FP64 index math)."""
        n = self._table.size
        ax = abs(float(x))
        if ax > self._halfwidth:
            return 0.0
        dx = self._halfwidth / (n - 1.0)
        one_dx = float(np.float32(1.0 / dx))
        axn = ax * one_dx
        i = min(int(math.floor(axn)), n - 2)
        t0 = float(self._table[i])
        dt = float(np.float32(self._table[i + 1] - self._table[i]))
        return float(np.float32(t0 + (axn - i) * dt))

    def _struct(self) -> _native.Kernel_t:
        return _native.Kernel_t(self._halfwidth, self._table.size, 0, self._table.ctypes.data)


# ------------------------------------------------------------------ ellipsoid
class Ellipsoid:
    """Reference ellipsoid (isce3::core::Ellipsoid), WGS84 by default."""

    def __init__(self, a: float = EarthSemiMajorAxis, e2: float = EarthEccentricitySquared):
        self.a = float(a)
        self.e2 = float(e2)

    @property
    def b(self) -> float:
        return self.a * math.sqrt(1.0 - self.e2)

    def lon_lat_to_xyz(self, llh) -> np.ndarray:
        llh = np.ascontiguousarray(llh, dtype=np.float64).reshape(3)
        out = np.empty(3)
        _native.lib().bp_lonlat_to_xyz(self.a, self.e2, _native.dptr(llh), _native.dptr(out))
        return out

    def xyz_to_lon_lat(self, xyz) -> np.ndarray:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(3)
        out = np.empty(3)
        _native.lib().bp_xyz_to_lonlat(self.a, self.e2, _native.dptr(xyz), _native.dptr(out))
        return out

    def r_east(self, lat: float) -> float:
        """Prime-vertical radius (core/Ellipsoid.h:117-120)."""
        return self.a / math.sqrt(1.0 - self.e2 * math.sin(lat) ** 2)

    def r_north(self, lat: float) -> float:
        """Meridional radius (core/Ellipsoid.h:127-130)."""
        return (self.a * (1.0 - self.e2)) / (1.0 - self.e2 * math.sin(lat) ** 2) ** 1.5

    def r_dir(self, hdg: float, lat: float) -> float:
        """Directional radius at heading `hdg` (clockwise from north,
        core/Ellipsoid.h:138-143)."""
        re, rn = self.r_east(lat), self.r_north(lat)
        return (re * rn) / (re * math.cos(hdg) ** 2 + rn * math.sin(hdg) ** 2)

    def n_vector(self, lon: float, lat: float) -> np.ndarray:
        """Outward unit normal at (lon, lat) (core/Ellipsoid.h:152-158)."""
        clat = math.cos(lat)
        return np.array([clat * math.cos(lon), clat * math.sin(lon), math.sin(lat)])


class ProjectionBase:
    """Map projection of an EPSG code (core/Projections.h, pybind
    core/Projections.cpp): ``forward(llh)`` from (lon, lat, h) in radians to
    projection coordinates, ``inverse(xyz)`` back; RuntimeError where the
    projection reports failure.  WGS84."""

    _name = "ProjectionBase"

    def __init__(self, code: int):
        self._code = int(code)
        probe = np.zeros(3)
        rc = _native.lib().bp_proj_forward(self._code, _native.dptr(probe), _native.dptr(probe.copy()))
        if rc < 0:
            raise errors.RuntimeError_(f"Unknown EPSG code (in factory): {self._code}")

    @property
    def code(self) -> int:
        return self._code

    @property
    def ellipsoid(self) -> "Ellipsoid":
        return Ellipsoid()

    def _call(self, fn, v, what):
        src = np.ascontiguousarray(v, dtype=np.float64).reshape(3)
        out = np.empty(3)
        rc = fn(self._code, _native.dptr(src), _native.dptr(out))
        if rc != 0:
            raise errors.RuntimeError_(f"{what} projection transformation failed")
        return out

    def forward(self, llh) -> np.ndarray:
        return self._call(_native.lib().bp_proj_forward, llh, "Forward")

    def inverse(self, xyz) -> np.ndarray:
        return self._call(_native.lib().bp_proj_inverse, xyz, "Inverse")

    def __repr__(self):
        return f"{self._name}(epsg={self._code})"


class LonLat(ProjectionBase):
    _name = "LonLat"

    def __init__(self):
        super().__init__(4326)


class Geocent(ProjectionBase):
    _name = "Geocent"

    def __init__(self):
        super().__init__(4978)


class UTM(ProjectionBase):
    """EPSG 32601-32660 (north zones 1-60), 32701-32760 (south)."""
    _name = "UTM"

    def __init__(self, code: int):
        code = int(code)
        if not (32600 < code <= 32660 or 32700 < code <= 32760):
            raise ValueError("In UTM::UTM - Invalid EPSG Code for UTM Projection. Received "
                             f"{code}, expected in ranges (32600,32660] or (32700,32760].")
        super().__init__(code)


class PolarStereo(ProjectionBase):
    """EPSG 3413 (north, Greenland) and 3031 (south, Antarctic)."""
    _name = "PolarStereo"

    def __init__(self, code: int):
        code = int(code)
        if code not in (3031, 3413):
            raise ValueError("In PolarStereo::PolarStereo - Invalid EPSG Code for Polar "
                             f"Stereographic projection. Received {code}, expected either "
                             "3031 (Antarctic) or 3413 (Greenland).")
        super().__init__(code)


class CEA(ProjectionBase):
    """Cylindrical equal area of EASE-Grid 2.0 (EPSG 6933)."""
    _name = "CEA"

    def __init__(self):
        super().__init__(6933)


def make_projection(epsg: int) -> ProjectionBase:
    """core/Projections.cpp:413-437 (createProj)."""
    epsg = int(epsg)
    if epsg == 4326:
        return LonLat()
    if epsg == 4978:
        return Geocent()
    if 32600 < epsg < 32800:
        return UTM(epsg)
    if epsg in (3031, 3413):
        return PolarStereo(epsg)
    if epsg == 6933:
        return CEA()
    raise errors.RuntimeError_(f"Unknown EPSG code (in factory): {epsg}")


class Basis:
    """Orthonormal basis x0, x1, x2 (core/Basis.h).  ``Basis(p, v)`` is the
    geocentric TCN frame: x2 = -p/|p|, x1 = x2 x v normalised,
    x0 = x1 x x2 normalised (core/Basis.h:50-58)."""

    def __init__(self, p=None, v=None):
        if p is None:
            self.x0, self.x1, self.x2 = np.eye(3)
            return
        p = np.asarray(p, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        n = -p / np.linalg.norm(p)
        c = np.cross(n, v)
        c = c / np.linalg.norm(c)
        t = np.cross(c, n)
        t = t / np.linalg.norm(t)
        self.x0, self.x1, self.x2 = t, c, n


def geodetic_tcn(x, v, ellipsoid: "Ellipsoid | None" = None) -> Basis:
    """TCN frame about the geodetic normal at x (core/Basis.cpp:21-26): the
    basis of (n_vector(lon, lat), v), so x2 points along the geodetic nadir."""
    ellipsoid = ellipsoid if ellipsoid is not None else Ellipsoid()
    lon, lat, _ = ellipsoid.xyz_to_lon_lat(x)
    return Basis(ellipsoid.n_vector(lon, lat), v)
