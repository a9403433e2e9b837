"""Drop-in ``backproject`` for isce3.focus.backproject / isce3.cuda.focus.backproject.

Signature, argument checks, error types and the return value follow
python/extensions/pybind_isce3/focus/Backproject.cpp:29-166 (CPU) and
python/extensions/pybind_isce3/cuda/focus/Backproject.cpp:25-116 (GPU, which
adds ``batch``).  The work runs in HIP kernels on the current GPU through
the C-ABI in include/isce3_amd_backproject.h.  Device tensors (torch, ROCm)
are used in place; pulses in host memory (numpy arrays, memory maps, CPU
tensors) are read only where the output needs them: the pixels' pulse window
is copied to the device in chunks of ``batch`` pulses, as the reference's GPU
path does (cxx/isce3/cuda/focus/Backproject.cu:641-688).
"""
from __future__ import annotations

import atexit
import ctypes as C
import threading
from typing import Optional

import numpy as np

from . import _native, errors
from .container import RadarGeometry
from .core import KernelF32, TabulatedKernelF32
from .geometry import DEMInterpolator

_TROPO = {"nodelay": 0, "tsx": 1}
_PRECISION = {"plain": _native.PRECISION_PLAIN, "split": _native.PRECISION_SPLIT}

DEFAULT_TOL_HEIGHT = 1e-5    # geometry/detail/Rdr2Geo.h:84
DEFAULT_TOL_AZTIME = 1e-7    # geometry/detail/Geo2Rdr.h:54


def parse_dry_tropo_model(s: str) -> int:
    """focus/DryTroposphereModel.cpp:20-33."""
    if s in _TROPO:
        return _TROPO[s]
    raise errors.InvalidArgument(
        f"expected one of {{'nodelay', 'tsx'}}, instead got '{s}'")


def parse_rdr2geo_params(params: dict) -> _native.Rdr2GeoParams_t:
    """pybind focus/Backproject.cpp:29-49."""
    out = _native.Rdr2GeoParams_t(DEFAULT_TOL_HEIGHT, 0.0, np.pi / 2)
    for key, val in dict(params).items():
        if key == "tol_height":
            out.tol_height = float(val)
        elif key == "look_min":
            out.look_min = float(val)
        elif key == "look_max":
            out.look_max = float(val)
        else:
            raise errors.InvalidArgument(f"unexpected rdr2geo_bracket keyword: {key}")
    return out


def parse_geo2rdr_params(params: dict) -> _native.Geo2RdrParams_t:
    """pybind focus/Backproject.cpp:52-78 (a None time means 'use the default')."""
    out = _native.Geo2RdrParams_t(DEFAULT_TOL_AZTIME, 0, 0, 0.0, 0.0)
    for key, val in dict(params).items():
        if key == "tol_aztime":
            out.tol_aztime = float(val)
        elif key == "time_start":
            if val is not None:
                out.have_time_start = 1
                out.time_start = float(val)
        elif key == "time_end":
            if val is not None:
                out.have_time_end = 1
                out.time_end = float(val)
        else:
            raise errors.InvalidArgument(f"unexpected geo2rdr_bracket keyword: {key}")
    return out


def geometry_struct(g: RadarGeometry) -> _native.RadarGeometry_t:
    grid = g.radar_grid
    gs = _native.Grid_t(grid.sensing_start, grid.prf, grid.starting_range,
                        grid.range_pixel_spacing, grid.length, grid.width, int(grid.lookside), 0)
    return _native.RadarGeometry_t(gs, g.orbit._struct(), g.doppler._struct())


class _Contexts:
    """One native context per HIP device, created on first use."""

    def __init__(self):
        self._lock = threading.Lock()
        self._ctx = {}

    def get(self, device: int):
        with self._lock:
            if device not in self._ctx:
                h = C.c_void_p()
                _native.check(_native.lib().bp_context_create(int(device), C.byref(h)))
                self._ctx[device] = h
            return self._ctx[device]

    def close(self):
        with self._lock:
            for h in self._ctx.values():
                _native.lib().bp_context_destroy(h)
            self._ctx.clear()


_contexts = _Contexts()
atexit.register(_contexts.close)


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _torch():
    import torch  # deferred: importing torch is slow and not needed for checks
    if not torch.cuda.is_available():
        raise errors.HipError("isce3_amd.focus.backproject needs a ROCm GPU (none visible)")
    return torch


def _shape_of(a):
    return tuple(int(s) for s in a.shape)


def backproject(out, out_geometry: RadarGeometry, signal_data, in_geometry: RadarGeometry,
                dem: DEMInterpolator, fc: float, ds: float, kernel: KernelF32,
                dry_tropo_model: str = "tsx", rdr2geo_params: Optional[dict] = None,
                geo2rdr_params: Optional[dict] = None, batch=1024, height=None, *,
                precision: str = "plain", options: Optional[dict] = None, stream=None,
                stats: Optional[dict] = None, max_window_bytes: int = 0) -> bool:
    """Focus in azimuth via time-domain backprojection.

    Parameters follow isce3.focus.backproject: ``out`` (complex64, written in
    place), ``out_geometry``, ``signal_data`` (range-compressed complex64),
    ``in_geometry``, ``dem``, ``fc`` (Hz), ``ds`` (desired azimuth resolution,
    m), ``kernel`` (TabulatedKernelF32), ``dry_tropo_model`` ("tsx" or
    "nodelay"), ``rdr2geo_params`` {tol_height, look_min, look_max},
    ``geo2rdr_params`` {tol_aztime, time_start, time_end}, ``batch`` (pulses
    per host-to-device copy when ``signal_data`` is in host memory; >= 1) and
    ``height`` (optional float32 output).  Arrays may be numpy arrays (memory
    maps included) or torch tensors; host pulses outside the output's pulse
    window are never read.

    Extensions: ``precision`` "plain" (reference semantics, FP64 inner loop)
    or "split" (FP64 per-pixel reference + FP32 deltas), ``options`` tiling
    dict {tile_rows, chunk, order, blocks_per_cu, staged, lane_rows}, ``stream`` (torch stream or raw hipStream_t int)
    ``stats`` (dict filled with updates, geometry_ms, sum_ms, and for host
    pulses h2d_bytes, window_pulses, bands, copy_ms) and ``max_window_bytes``
    (device bytes for the pulse window of host pulses, 0 = half the free
    memory; a larger window is focused in bands of output lines).

    Returns True if any pixel failed to converge (those pixels are NaN).
    """
    # The CPU binding takes height as the 12th positional argument.
    if batch is not None and not isinstance(batch, (int, np.integer)) and height is None:
        height, batch = batch, 1024
    if out.ndim != 2:
        raise errors.InvalidArgument("output array must be 2-D")
    if _shape_of(out) != (out_geometry.grid_length, out_geometry.grid_width):
        raise errors.InvalidArgument("output array shape must match output radar grid shape")
    if signal_data.ndim != 2:
        raise errors.InvalidArgument("input signal data must be 2-D")
    if _shape_of(signal_data) != (in_geometry.grid_length, in_geometry.grid_width):
        raise errors.InvalidArgument("input signal data shape must match input radar grid shape")
    if height is not None and _shape_of(height) != (out_geometry.grid_length,
                                                    out_geometry.grid_width):
        raise errors.InvalidArgument("height array shape must match output radar grid shape")
    tropo = parse_dry_tropo_model(dry_tropo_model)
    r2g = parse_rdr2geo_params(rdr2geo_params or {})
    g2r = parse_geo2rdr_params(geo2rdr_params or {})
    if batch is None or int(batch) < 1:
        raise errors.DomainError("batch size must be > 0")
    if out_geometry.reference_epoch != in_geometry.reference_epoch:
        raise errors.RuntimeError_("input reference epoch must match output reference epoch")
    if not isinstance(kernel, TabulatedKernelF32):
        raise errors.InvalidArgument("kernel must be a TabulatedKernelF32")
    if precision not in _PRECISION:
        raise errors.InvalidArgument(f"precision must be 'plain' or 'split', got {precision!r}")
    _check_out_array(out, "output")
    if height is not None:
        _check_out_array(height, "height", complex_=False)

    torch = _torch()
    device = _device_of(torch, out, signal_data, height)
    with torch.cuda.device(device):
        return _run(torch, device, out, out_geometry, signal_data, in_geometry, dem, fc, ds,
                    kernel, tropo, r2g, g2r, height, _PRECISION[precision], options, stream,
                    stats, int(batch), int(max_window_bytes))


def _check_out_array(a, what, complex_=True):
    if _is_torch(a):
        import torch
        want = torch.complex64 if complex_ else torch.float32
        if a.dtype != want or not a.is_contiguous():
            raise errors.InvalidArgument(f"{what} must be a contiguous {want} tensor")
        return
    a = np.asarray(a) if not isinstance(a, np.ndarray) else a
    want = np.complex64 if complex_ else np.float32
    if a.dtype != want or not a.flags.c_contiguous or not a.flags.writeable:
        raise errors.InvalidArgument(
            f"{what} must be a writeable C-contiguous {np.dtype(want).name} array")


def _device_of(torch, *arrays) -> int:
    for a in arrays:
        if a is not None and _is_torch(a) and a.is_cuda:
            return a.device.index if a.device.index is not None else torch.cuda.current_device()
    return torch.cuda.current_device()


def _stream_handle(torch, stream):
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _device_input(torch, device, sig):
    if _is_torch(sig):
        d_in = sig.contiguous() if sig.is_cuda else sig.contiguous().to(f"cuda:{device}")
        if d_in.dtype != torch.complex64:
            d_in = d_in.to(torch.complex64)
        return d_in
    h = np.ascontiguousarray(sig, dtype=np.complex64)
    if not h.flags.writeable:  # torch tensors over read-only numpy memory are undefined
        h = h.copy()
    return torch.from_numpy(h).to(f"cuda:{device}", non_blocking=False)


def _host_pulses(torch, sig) -> np.ndarray:
    """C-contiguous complex64 view of pulses in host memory (numpy array,
    memory map or CPU tensor); copied only when the layout or type differs."""
    if _is_torch(sig):
        sig = sig.contiguous()
        if sig.dtype != torch.complex64:
            sig = sig.to(torch.complex64)
        return sig.numpy()
    if isinstance(sig, np.ndarray) and sig.dtype == np.complex64 and sig.flags.c_contiguous:
        return sig
    return np.ascontiguousarray(sig, dtype=np.complex64)


def _device_output(torch, device, a, dtype):
    """Device buffer written by the kernels: the array itself when it is a
    device tensor, else a staging tensor copied back by _copy_back."""
    if a is None:
        return None
    if _is_torch(a) and a.is_cuda:
        return a
    return torch.empty(_shape_of(a), dtype=dtype, device=f"cuda:{device}")


def _copy_back(torch, a, d):
    if a is None or d is a:
        return
    res = d.cpu().numpy()
    if _is_torch(a):
        a.copy_(torch.from_numpy(res))
    else:
        a[...] = res


def _options_struct(options) -> _native.Options_t:
    opt = _native.Options_t()
    for k, v in (options or {}).items():
        if k not in ("tile_rows", "chunk", "fit_degree", "order", "blocks_per_cu", "staged",
                     "lane_rows"):
            raise errors.InvalidArgument(f"unknown option {k!r}")
        setattr(opt, k, int(v))
    return opt


def _fill_stats(stats, st):
    if stats is not None:
        stats.update(updates=int(st.updates), geometry_ms=float(st.geometry_ms),
                     sum_ms=float(st.sum_ms), step_ms=float(st.step_ms),
                     any_failed=bool(st.any_failed))


def _run(torch, device, out, ogeom, sig, igeom, dem, fc, ds, kernel, tropo, r2g, g2r, height,
         precision, options, stream, stats, batch=1024, max_window_bytes=0):
    if not (_is_torch(sig) and sig.is_cuda):
        return _run_host(torch, device, out, ogeom, sig, igeom, dem, fc, ds, kernel, tropo, r2g,
                         g2r, height, precision, options, stream, stats, batch,
                         max_window_bytes)
    hstream = _stream_handle(torch, stream)
    d_in = _device_input(torch, device, sig)
    d_out = _device_output(torch, device, out, torch.complex64)
    d_h = _device_output(torch, device, height, torch.float32)
    og = geometry_struct(ogeom)
    ig = geometry_struct(igeom)
    dm = dem._struct()
    kr = kernel._struct()
    opt = _options_struct(options)
    st = _native.Stats_t()
    failed = C.c_int(0)
    ctx = _contexts.get(device)
    # keep the host arrays referenced by the structs alive for the call
    _keep = (ogeom, igeom, dem, kernel)
    rc = _native.lib().bp_backproject(
        ctx, C.byref(og), C.byref(ig), C.byref(dm), float(fc), float(ds), C.byref(kr),
        int(tropo), C.byref(r2g), C.byref(g2r), C.c_void_p(d_in.data_ptr()),
        C.c_void_p(d_out.data_ptr()), C.c_void_p(d_h.data_ptr()) if d_h is not None else None,
        int(precision), C.byref(opt), C.c_void_p(hstream), C.byref(failed), C.byref(st))
    del _keep
    _native.check(rc)
    _copy_back(torch, out, d_out)
    _copy_back(torch, height, d_h)
    _fill_stats(stats, st)
    return bool(failed.value)


def _run_host(torch, device, out, ogeom, sig, igeom, dem, fc, ds, kernel, tropo, r2g, g2r,
              height, precision, options, stream, stats, batch, max_window_bytes):
    """Pulses in host memory: bp_backproject_host copies the pixels' pulse
    window only (in chunks of ``batch`` pulses)."""
    hstream = _stream_handle(torch, stream)
    h_in = _host_pulses(torch, sig)
    d_out = _device_output(torch, device, out, torch.complex64)
    d_h = _device_output(torch, device, height, torch.float32)
    og = geometry_struct(ogeom)
    ig = geometry_struct(igeom)
    dm = dem._struct()
    kr = kernel._struct()
    opt = _options_struct(options)
    st = _native.Stats_t()
    hst = _native.HostStats_t()
    failed = C.c_int(0)
    ctx = _contexts.get(device)
    _keep = (ogeom, igeom, dem, kernel, h_in)
    rc = _native.lib().bp_backproject_host(
        ctx, C.byref(og), C.byref(ig), C.byref(dm), float(fc), float(ds), C.byref(kr),
        int(tropo), C.byref(r2g), C.byref(g2r), C.c_void_p(h_in.ctypes.data), int(batch),
        int(max_window_bytes), C.c_void_p(d_out.data_ptr()),
        C.c_void_p(d_h.data_ptr()) if d_h is not None else None, int(precision), C.byref(opt),
        C.c_void_p(hstream), C.byref(failed), C.byref(st), C.byref(hst))
    del _keep
    _native.check(rc)
    _copy_back(torch, out, d_out)
    _copy_back(torch, height, d_h)
    _fill_stats(stats, st)
    if stats is not None:
        stats.update(h2d_bytes=int(hst.h2d_bytes), window_pulses=int(hst.window_pulses),
                     bands=int(hst.bands), copy_ms=float(hst.copy_ms))
    return bool(failed.value)


def _same_orbit(a, b) -> bool:
    if a is b:
        return True
    sa, sb = a._struct(), b._struct()
    return (sa.t0 == sb.t0 and sa.dt == sb.dt and sa.n == sb.n and sa.method == sb.method
            and a.reference_epoch == b.reference_epoch
            and np.array_equal(a._pos, b._pos) and np.array_equal(a._vel, b._vel))


def _same_lut(a, b) -> bool:
    if a is b:
        return True
    sa, sb = a._struct(), b._struct()
    same = all(getattr(sa, f) == getattr(sb, f) for f, _ in _native.Lut2d_t._fields_
               if f != "data")
    return same and (not a.have_data or np.array_equal(a.data, b.data))


def backproject_blocks(outs, out_geometries, signal_data, in_geometry: RadarGeometry,
                       dem: DEMInterpolator, fc: float, ds: float, kernel: KernelF32,
                       dry_tropo_model: str = "tsx", rdr2geo_params: Optional[dict] = None,
                       geo2rdr_params: Optional[dict] = None, heights=None, *,
                       precision: str = "plain", options: Optional[dict] = None, stream=None,
                       stats: Optional[dict] = None):
    """Focus several output blocks of one raw data set in one pass.

    Equivalent to calling :func:`backproject` once per block, as the RSLC
    workflow does over its processing blocks
    (python/packages/nisar/workflows/focus.py:1986-2004), but the pixels of
    all blocks go through one geometry launch and one coherent-sum launch.
    ``outs`` and ``out_geometries`` are equal-length lists; every output
    geometry must share one orbit and one Doppler LUT.  ``heights`` is None
    or a list (entries may be None).  Returns a list of per-block flags, True
    where some pixel of that block failed to converge.
    """
    outs = list(outs)
    out_geometries = list(out_geometries)
    if len(outs) != len(out_geometries):
        raise errors.InvalidArgument("need one output geometry per output array")
    heights = [None] * len(outs) if heights is None else list(heights)
    if len(heights) != len(outs):
        raise errors.InvalidArgument("need one height entry per output array")
    if signal_data.ndim != 2:
        raise errors.InvalidArgument("input signal data must be 2-D")
    if _shape_of(signal_data) != (in_geometry.grid_length, in_geometry.grid_width):
        raise errors.InvalidArgument("input signal data shape must match input radar grid shape")
    for out, og, h in zip(outs, out_geometries, heights):
        if out.ndim != 2 or _shape_of(out) != (og.grid_length, og.grid_width):
            raise errors.InvalidArgument("output array shape must match output radar grid shape")
        if h is not None and _shape_of(h) != (og.grid_length, og.grid_width):
            raise errors.InvalidArgument("height array shape must match output radar grid shape")
        if og.reference_epoch != in_geometry.reference_epoch:
            raise errors.RuntimeError_("input reference epoch must match output reference epoch")
        _check_out_array(out, "output")
        if h is not None:
            _check_out_array(h, "height", complex_=False)
    if out_geometries:
        o0, d0 = out_geometries[0].orbit, out_geometries[0].doppler
        for og in out_geometries[1:]:
            if not _same_orbit(og.orbit, o0) or not _same_lut(og.doppler, d0):
                raise errors.InvalidArgument(
                    "all output blocks must share one orbit and one Doppler LUT")
    tropo = parse_dry_tropo_model(dry_tropo_model)
    r2g = parse_rdr2geo_params(rdr2geo_params or {})
    g2r = parse_geo2rdr_params(geo2rdr_params or {})
    if not isinstance(kernel, TabulatedKernelF32):
        raise errors.InvalidArgument("kernel must be a TabulatedKernelF32")
    if precision not in _PRECISION:
        raise errors.InvalidArgument(f"precision must be 'plain' or 'split', got {precision!r}")
    if not outs:
        return []

    torch = _torch()
    device = _device_of(torch, signal_data, *outs, *heights)
    with torch.cuda.device(device):
        hstream = _stream_handle(torch, stream)
        d_in = _device_input(torch, device, signal_data)
        d_outs = [_device_output(torch, device, o, torch.complex64) for o in outs]
        d_hs = [_device_output(torch, device, h, torch.float32) for h in heights]
        blocks = (_native.OutBlock_t * len(outs))()
        for i, og in enumerate(out_geometries):
            blocks[i].grid = geometry_struct(og).grid
            blocks[i].out_dev = d_outs[i].data_ptr()
            blocks[i].height_dev = d_hs[i].data_ptr() if d_hs[i] is not None else None
        g0 = out_geometries[0]
        oorb = g0.orbit._struct()
        odop = g0.doppler._struct()
        ig = geometry_struct(in_geometry)
        dm = dem._struct()
        kr = kernel._struct()
        opt = _options_struct(options)
        st = _native.Stats_t()
        failed = C.c_int(0)
        ctx = _contexts.get(device)
        _keep = (out_geometries, in_geometry, dem, kernel)
        rc = _native.lib().bp_backproject_blocks(
            ctx, C.byref(oorb), C.byref(odop), blocks, len(outs), C.byref(ig), C.byref(dm),
            float(fc), float(ds), C.byref(kr), int(tropo), C.byref(r2g), C.byref(g2r),
            C.c_void_p(d_in.data_ptr()), _PRECISION[precision], C.byref(opt),
            C.c_void_p(hstream), C.byref(failed), C.byref(st))
        del _keep
        _native.check(rc)
        for o, d in zip(outs, d_outs):
            _copy_back(torch, o, d)
        for h, d in zip(heights, d_hs):
            _copy_back(torch, h, d)
        _fill_stats(stats, st)
        return [bool(blocks[i].failed) for i in range(len(outs))]


class Scene:
    """One raw data set resident on the GPU, focused tile by tile through a
    captured HIP-graph step (bp_scene_* in include/isce3_amd_backproject.h).

    Holds the arguments that the per-block isce3.focus.backproject calls of
    the RSLC workflow share (python/packages/nisar/workflows/focus.py:1986-2004):
    the range-compressed pulses ``signal_data`` (a ROCm complex64 tensor that
    must stay alive and unchanged), ``in_geometry``, the output ``out_orbit``
    and ``out_doppler``, ``dem``, ``fc``, ``ds``, ``kernel``, the troposphere
    model and solver parameters.  Orbits, Doppler tables, DEM and kernel
    tables are uploaded once; per-pixel scratch is reserved for up to
    ``max_pixels`` pixels in up to ``max_blocks`` output blocks per call.
    :meth:`focus` gives, block for block, the result of
    :func:`backproject_blocks` with the same arguments.  One host thread at a
    time may use a scene.
    """

    def __init__(self, signal_data, in_geometry: RadarGeometry, out_orbit, out_doppler,
                 dem: DEMInterpolator, fc: float, ds: float, kernel: KernelF32,
                 dry_tropo_model: str = "tsx", rdr2geo_params: Optional[dict] = None,
                 geo2rdr_params: Optional[dict] = None, *, precision: str = "split",
                 max_pixels: int = 1 << 22, max_blocks: int = 64):
        self._h = None
        if precision not in _PRECISION:
            raise errors.InvalidArgument(f"precision must be 'plain' or 'split', got {precision!r}")
        if not isinstance(kernel, TabulatedKernelF32):
            raise errors.InvalidArgument("kernel must be a TabulatedKernelF32")
        if not _is_torch(signal_data) or not signal_data.is_cuda:
            raise errors.InvalidArgument("a scene's signal data must be a device tensor")
        if signal_data.ndim != 2 or \
                _shape_of(signal_data) != (in_geometry.grid_length, in_geometry.grid_width):
            raise errors.InvalidArgument("input signal data shape must match input radar grid shape")
        if out_orbit.reference_epoch != in_geometry.reference_epoch:
            raise errors.RuntimeError_("input reference epoch must match output reference epoch")
        torch = _torch()
        if signal_data.dtype != torch.complex64 or not signal_data.is_contiguous():
            raise errors.InvalidArgument("signal data must be a C-contiguous complex64 tensor")
        self.device = signal_data.device.index or 0
        self.precision = precision
        self.max_pixels = int(max_pixels)
        self.max_blocks = int(max_blocks)
        self.epoch = out_orbit.reference_epoch
        self._signal = signal_data
        oorb = out_orbit._struct()
        odop = out_doppler._struct()
        ig = geometry_struct(in_geometry)
        dm = dem._struct()
        kr = kernel._struct()
        r2g = parse_rdr2geo_params(rdr2geo_params or {})
        g2r = parse_geo2rdr_params(geo2rdr_params or {})
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            ctx = _contexts.get(self.device)
            _keep = (in_geometry, out_orbit, out_doppler, dem, kernel)
            rc = _native.lib().bp_scene_create(
                ctx, C.byref(oorb), C.byref(odop), C.byref(ig), C.byref(dm), float(fc),
                float(ds), C.byref(kr), parse_dry_tropo_model(dry_tropo_model), C.byref(r2g),
                C.byref(g2r), C.c_void_p(signal_data.data_ptr()), _PRECISION[precision],
                self.max_pixels, self.max_blocks, C.byref(h))
            del _keep
        _native.check(rc)
        self._h = h

    def _grid_struct(self, grid) -> _native.Grid_t:
        t0 = grid.sensing_start
        if grid.ref_epoch != self.epoch:
            t0 = t0 + (grid.ref_epoch - self.epoch).total_seconds()
        return _native.Grid_t(t0, grid.prf, grid.starting_range, grid.range_pixel_spacing,
                              grid.length, grid.width, int(grid.lookside), 0)

    def focus(self, outs, grids, heights=None, *, options: Optional[dict] = None, stream=None,
              stats: Optional[dict] = None):
        """Focus output blocks: ``outs`` are complex64 device tensors of the
        shapes of ``grids`` (RadarGridParameters), ``heights`` None or a list
        of float32 device tensors / None.  Returns per-block failure flags."""
        if self._h is None:
            raise errors.RuntimeError_("scene is closed")
        outs, grids = list(outs), list(grids)
        if len(outs) != len(grids):
            raise errors.InvalidArgument("need one output grid per output array")
        heights = [None] * len(outs) if heights is None else list(heights)
        if len(heights) != len(outs):
            raise errors.InvalidArgument("need one height entry per output array")
        torch = _torch()
        blocks = (_native.OutBlock_t * max(len(outs), 1))()
        for i, (o, g, hh) in enumerate(zip(outs, grids, heights)):
            if not _is_torch(o) or not o.is_cuda or o.dtype != torch.complex64 \
                    or not o.is_contiguous() or _shape_of(o) != (g.length, g.width):
                raise errors.InvalidArgument(
                    "outputs must be C-contiguous complex64 device tensors of the grid shape")
            if hh is not None and (not _is_torch(hh) or hh.dtype != torch.float32
                                   or not hh.is_contiguous() or _shape_of(hh) != (g.length, g.width)):
                raise errors.InvalidArgument(
                    "heights must be C-contiguous float32 device tensors of the grid shape")
            blocks[i].grid = self._grid_struct(g)
            blocks[i].out_dev = o.data_ptr()
            blocks[i].height_dev = hh.data_ptr() if hh is not None else None
        opt = _options_struct(options)
        st = _native.Stats_t()
        failed = C.c_int(0)
        with torch.cuda.device(self.device):
            hstream = _stream_handle(torch, stream)
            rc = _native.lib().bp_scene_focus(self._h, blocks, len(outs), C.byref(opt),
                                              C.c_void_p(hstream), C.byref(failed), C.byref(st))
        _native.check(rc)
        _fill_stats(stats, st)
        return [bool(blocks[i].failed) for i in range(len(outs))]

    def close(self):
        if self._h is not None:
            _native.lib().bp_scene_destroy(self._h)
            self._h = None
            self._signal = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class TileServer:
    """Continuous batching of tile requests against one :class:`Scene`
    (bp_server_* / bp_serve_* in include/isce3_amd_backproject.h): a native
    worker thread focuses every pending request that fits the scene in one
    step and starts on the requests that arrived meanwhile.  ``submit``
    returns a ticket at once; ``poll`` / ``wait`` report its completion (the
    per-request failure flag).  The server owns the scene until
    :meth:`close`, which finishes every accepted request first."""

    def __init__(self, scene: Scene, options: Optional[dict] = None, max_pending: int = 1024):
        self._h = None
        if scene._h is None:
            raise errors.RuntimeError_("scene is closed")
        self.scene = scene
        opt = _options_struct(options)
        h = C.c_void_p()
        _native.check(_native.lib().bp_server_create(scene._h, C.byref(opt), int(max_pending),
                                                     C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, out, grid, height=None) -> int:
        """Queue one output block (``out`` a complex64 device tensor of the
        grid's shape, kept alive until the ticket completes)."""
        if self._h is None:
            raise errors.RuntimeError_("server is closed")
        torch = _torch()
        if not _is_torch(out) or not out.is_cuda or out.dtype != torch.complex64 \
                or not out.is_contiguous() or _shape_of(out) != (grid.length, grid.width):
            raise errors.InvalidArgument(
                "outputs must be C-contiguous complex64 device tensors of the grid shape")
        if height is not None and (not _is_torch(height) or height.dtype != torch.float32 or
                                   not height.is_contiguous() or
                                   _shape_of(height) != (grid.length, grid.width)):
            raise errors.InvalidArgument(
                "heights must be C-contiguous float32 device tensors of the grid shape")
        blk = _native.OutBlock_t()
        blk.grid = self.scene._grid_struct(grid)
        blk.out_dev = out.data_ptr()
        blk.height_dev = height.data_ptr() if height is not None else None
        # the worker runs on its own stream: work queued on the caller's
        # stream (e.g. the buffers' allocation fills) completes first
        with torch.cuda.device(self.scene.device):
            torch.cuda.current_stream().synchronize()
        t = C.c_uint64()
        _native.check(_native.lib().bp_serve_submit(self._h, C.byref(blk), C.byref(t)))
        self._keep[t.value] = (out, height)
        return int(t.value)

    def poll(self, ticket: int):
        """None while pending, else the request's failure flag."""
        done, failed = C.c_int(0), C.c_int(0)
        _native.check(_native.lib().bp_serve_poll(self._h, C.c_uint64(ticket), C.byref(done),
                                                  C.byref(failed)))
        if not done.value:
            return None
        self._keep.pop(ticket, None)
        return bool(failed.value)

    def wait(self, ticket: int) -> bool:
        """Block until the request is focused; returns its failure flag."""
        failed = C.c_int(0)
        _native.check(_native.lib().bp_serve_wait(self._h, C.c_uint64(ticket), C.byref(failed)))
        self._keep.pop(ticket, None)
        return bool(failed.value)

    def close(self):
        if self._h is not None:
            _native.lib().bp_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


# range compression, the step before backprojection (isce3.focus.RangeComp,
# isce3.focus.form_linear_chirp)
from .rangecomp import RangeComp, form_linear_chirp  # noqa: E402,F401
from .planning import get_radar_velocities, get_sar_duration  # noqa: E402,F401
from .presum import (apply_presum_weights, compute_ids_from_mask, fill_weights,  # noqa: E402,F401
                     get_presum_weights, get_unique_ids)
