"""Presumming: BLU reconstruction of non-uniformly sampled (dithered-PRF,
gapped) pulses onto a uniform pulse grid.

Look-alikes of ``isce3.focus.get_presum_weights`` / ``fill_weights`` /
``apply_presum_weights`` / ``compute_ids_from_mask`` / ``get_unique_ids``
(pybind focus/Presum.cpp:26-220, cxx/isce3/focus/Presum.{h,icc,cpp}) and of
the workflow's ``nisar.workflows.focus.resample`` (focus.py:973-1062).  The
weighted, Doppler-deramped pulse sums run in HIP kernels behind
include/isce3_amd_presum.h; the BLU weight systems (a few per output pulse,
at most 63 unknowns) are solved in the library's host code.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native, errors


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _torch(device: int = 0):
    import torch
    if not torch.cuda.is_available():
        raise errors.HipError("presumming needs a ROCm GPU (none visible)")
    return torch


def get_presum_weights(acorr, t, tout: float):
    """(offset, weights): weights.dot(x[offset:offset + len(weights)])
    reconstructs the sample at ``tout`` from samples at sorted times ``t``
    (Presum.icc:9-60).  ``acorr`` is an AzimuthKernel or BartlettKernel."""
    if not hasattr(acorr, "_struct"):
        raise errors.InvalidArgument("acorr must be an AzimuthKernel or BartlettKernel")
    t = np.ascontiguousarray(t, dtype=np.float64).ravel()
    w = np.empty(max(len(t), 1))
    offset = C.c_int64()
    nw = C.c_int64()
    _native.check(_native.lib().bp_presum_weights(
        C.byref(acorr._struct()), t.ctypes.data, len(t), float(tout), C.byref(offset),
        w.ctypes.data, len(w), C.byref(nw)))
    return int(offset.value), w[:nw.value].copy()


def fill_weights(ids, lut: dict) -> np.ndarray:
    """Matrix (nw, len(ids)) whose column i is ``lut[ids[i]]``
    (Presum.cpp:5-30)."""
    if not lut:
        raise errors.LengthError("Need at least one weight vector in LUT.")
    keys = list(lut.keys())
    vecs = [np.asarray(lut[k], dtype=np.float64).ravel() for k in keys]
    nw = len(vecs[0])
    if any(len(v) != nw for v in vecs):
        raise errors.LengthError("Weight vector lengths must all match.")
    table = np.stack(vecs, axis=1) if nw else np.zeros((0, len(keys)))
    index = {k: i for i, k in enumerate(keys)}
    try:
        cols = np.fromiter((index[int(i)] for i in np.asarray(ids).ravel()), dtype=np.int64)
    except KeyError as e:
        raise IndexError(f"id {e.args[0]} not in LUT") from None
    return np.ascontiguousarray(table[:, cols])


def compute_ids_from_mask(mask) -> np.ndarray:
    """ids[i] with bit j set where mask[i, j] (pybind Presum.cpp:96-125);
    at most 63 columns."""
    m = np.ascontiguousarray(mask, dtype=np.bool_)
    if m.ndim != 2:
        raise errors.LengthError("Expected 2D mask")
    ids = np.empty(m.shape[0], np.int64)
    _native.check(_native.lib().bp_presum_compute_ids(m.ctypes.data, m.shape[0], m.shape[1],
                                                      ids.ctypes.data))
    return ids


def get_unique_ids(ids) -> np.ndarray:
    """Sorted distinct ids (pybind Presum.cpp:70-94; same result as np.unique)."""
    return np.unique(np.asarray(ids, dtype=np.int64))


def apply_presum_weights(out, pulse_times, doppler, weights, pulses, device: int = 0) -> None:
    """out[j] = sum_i float(weights[i, j]) exp(-2 pi i pulse_times[i] doppler[j])
    pulses[i, j] in complex64 (pybind Presum.cpp:26-66), on the GPU.  Arrays
    may be numpy or ROCm torch tensors; ``out`` is written in place."""
    pt = np.asarray(pulse_times) if not _is_torch(pulse_times) else pulse_times
    fd = np.asarray(doppler) if not _is_torch(doppler) else doppler
    if pt.ndim != 1 or fd.ndim != 1:
        raise errors.LengthError("Expected 1D pulse_times and doppler arrays.")
    m, n = int(pt.shape[0]), int(fd.shape[0])
    if tuple(out.shape) != (n,) and int(np.prod(out.shape)) != n:
        raise errors.LengthError("Output vector length doesn't match Doppler")
    if weights.ndim != 2 or tuple(weights.shape) != (m, n):
        raise errors.LengthError("Weight dimensions don't match time and Doppler vectors")
    if pulses.ndim != 2 or tuple(pulses.shape) != (m, n):
        raise errors.LengthError("Raw data dimensions don't match time and Doppler vectors")
    torch = _torch(device)
    dev = f"cuda:{device}"

    def up(a, dtype):
        if _is_torch(a):
            return a.to(device=dev, dtype=dtype).contiguous()
        return torch.from_numpy(np.ascontiguousarray(a, dtype={torch.float64: np.float64,
                                                               torch.complex64: np.complex64}[dtype])).to(dev)

    with torch.cuda.device(device):
        d_t, d_fd = up(pt, torch.float64), up(fd, torch.float64)
        d_w, d_x = up(weights, torch.float64), up(pulses, torch.complex64)
        in_place = (_is_torch(out) and out.is_cuda and out.dtype == torch.complex64 and
                    out.is_contiguous())
        d_out = out if in_place else torch.empty((n,), dtype=torch.complex64, device=dev)
        ptr = lambda a: C.c_void_p(a.data_ptr())  # noqa: E731
        _native.check(_native.lib().bp_presum_apply_weights(
            ptr(d_out), ptr(d_t), ptr(d_fd), ptr(d_w), ptr(d_x), m, n,
            C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        if not in_place:
            res = d_out.cpu().numpy().reshape(out.shape)
            if _is_torch(out):
                out.copy_(torch.from_numpy(res))
            else:
                out[...] = res


def resample(raw, t, grid, swaths, orbit, doppler, L: float = 12.0, fn=None, device: int = 0):
    """Fill gaps and resample raw pulses to the uniform grid ``grid`` by BLU
    reconstruction (nisar.workflows.focus.resample, focus.py:973-1062):
    output pulse i at t[0] + i / grid.prf, the azimuth autocorrelation scaled
    by L / |v(tout)|, weights per pattern of valid samples (``swaths``:
    int array (ns, nt, 2) of [start, stop) valid range bins per pulse),
    Doppler deramp at ``doppler.eval(tout, r)``.

    ``raw`` (grid.length x grid.width complex64) may be numpy or a ROCm torch
    tensor.  Returns a device tensor for device input, else a numpy array (a
    memory map of file ``fn`` when given)."""
    assert tuple(raw.shape) == (grid.length, grid.width)
    assert len(t) == raw.shape[0]
    assert grid.ref_epoch == orbit.reference_epoch
    t = np.ascontiguousarray(t, dtype=np.float64)
    sw = np.ascontiguousarray(swaths, dtype=np.int64)
    if sw.ndim != 3 or sw.shape[1] != len(t) or sw.shape[2] != 2:
        raise errors.LengthError("swaths must have shape (ns, nt, 2)")
    sw32 = np.ascontiguousarray(np.clip(sw, np.iinfo(np.int32).min, np.iinfo(np.int32).max)
                                .astype(np.int32))
    torch = _torch(device)
    dev = f"cuda:{device}"
    from ._native import Grid_t
    g = Grid_t(grid.sensing_start, grid.prf, grid.starting_range, grid.range_pixel_spacing,
               grid.length, grid.width, int(grid.lookside), 0)
    o = orbit._struct()
    lut = doppler._struct()
    with torch.cuda.device(device):
        d_raw = (raw.to(device=dev, dtype=torch.complex64).contiguous() if _is_torch(raw)
                 else torch.from_numpy(np.ascontiguousarray(raw, dtype=np.complex64)).to(dev))
        d_out = torch.empty((grid.length, grid.width), dtype=torch.complex64, device=dev)
        _native.check(_native.lib().bp_presum_resample(
            int(device), t.ctypes.data, len(t), sw32.ctypes.data, sw32.shape[0], C.byref(g),
            C.byref(o), C.byref(lut), float(L), C.c_void_p(d_raw.data_ptr()),
            C.c_void_p(d_out.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        if _is_torch(raw):
            return d_out
        res = d_out.cpu().numpy()
    if fn is None:
        return res
    mm = np.memmap(fn, mode="w+", shape=res.shape, dtype=np.complex64)
    mm[...] = res
    return mm
