"""Finishing focused RSLC blocks on the GPU: statistics, range correction,
accumulation and storage encoding.

Look-alikes of
  * ``isce3.math.StatsFloat32`` / ``StatsRealImagFloat32`` (cxx/isce3/math/
    Stats.h:15-77, Stats.cpp:33-165, pybind math/Stats.cpp): records with the
    same fields, ``update`` by Chan's pairwise rule and the same ``str`` form;
    the statistics of an array are computed by the HIP finisher;
  * ``isce3.core.types.complex32`` / ``to_complex32`` / ``truncate_mantissa``
    (python/packages/isce3/core/types.py:5, 38-42, 116-171), run on the GPU;
  * ``nisar.workflows.focus.BackgroundWriter`` (focus.py:817-914): range
    correction, mixed-mode accumulation, block statistics and encoding in one
    HIP pass per block (include/isce3_amd_rslc.h), the dataset write on a
    worker thread.  The dataset is any object with ``dtype``,
    ``write_direct(array, dest_sel=...)`` and ``__getitem__`` (an
    ``h5py.Dataset`` in the workflow).
"""
from __future__ import annotations

import ctypes as C
import logging
import math
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _native, errors

log = logging.getLogger("isce3_amd.rslc")

complex32 = np.dtype([("r", np.float16), ("i", np.float16)])

DATA_TYPES = {"complex64": 0, "complex32": 1, "complex64_zero_mantissa": 2}


# ---------------------------------------------------------------- statistics

class StatsFloat32:
    """Stats<float>: min, max, mean and sample standard deviation of the
    non-NaN values seen so far (Stats.h:15-49)."""

    _name = "StatsFloat32"

    def __init__(self, x=None):
        self.min = math.nan
        self.max = math.nan
        self.mean = 0.0
        self.n_valid = 0
        self._mean64 = 0.0
        self._m2 = 0.0
        if x is not None:
            self.update(x)

    @classmethod
    def _from_struct(cls, s: _native.StatsF32_t) -> "StatsFloat32":
        o = cls()
        o.min, o.max, o.mean = float(s.min), float(s.max), float(s.mean)
        o.n_valid = int(s.n_valid)
        o._mean64, o._m2 = float(s.real_valued_mean), float(s.square_diff_sum)
        return o

    @property
    def sample_stddev(self) -> float:
        if self.n_valid > 1:
            return float(np.float32(math.sqrt(self._m2 / (self.n_valid - 1))))
        return 0.0

    def _merge(self, other: "StatsFloat32") -> None:
        # Stats.cpp:64-96 (Chan et al.)
        if other.n_valid <= 0:
            return
        nself, ntotal = self.n_valid, self.n_valid + other.n_valid
        self.n_valid = ntotal
        delta = other._mean64 - self._mean64
        self._mean64 += delta * other.n_valid / ntotal
        self._m2 += other._m2 + delta * delta * nself * other.n_valid / ntotal
        self.mean = float(np.float32(self._mean64))
        if math.isnan(self.max) or other.max > self.max:
            self.max = other.max
        if math.isnan(self.min) or other.min < self.min:
            self.min = other.min

    def update(self, x) -> None:
        """Merge another StatsFloat32, or the statistics of a real float32
        array (computed on the GPU)."""
        if isinstance(x, StatsFloat32):
            self._merge(x)
            return
        a = np.asarray(x, dtype=np.float32)
        z = np.empty(a.shape, np.complex64)
        z.real = a
        z.imag = 0.0
        self._merge(_array_stats(z).real)

    def __str__(self) -> str:
        return (f"{self._name}(n_valid={self.n_valid}, mean={self.mean}, min={self.min}, "
                f"max={self.max}, sample_stddev={self.sample_stddev})")

    __repr__ = __str__


class StatsRealImagFloat32:
    """StatsRealImag<float>: separate statistics of the real and imaginary
    parts; a sample counts when neither part is NaN (Stats.h:54-77)."""

    def __init__(self, z=None):
        self.real = StatsFloat32()
        self.imag = StatsFloat32()
        self.n_valid = 0
        if z is not None:
            self.update(z)

    @classmethod
    def _from_struct(cls, s: _native.StatsRealImagF32_t) -> "StatsRealImagFloat32":
        o = cls()
        o.real = StatsFloat32._from_struct(s.real)
        o.imag = StatsFloat32._from_struct(s.imag)
        o.n_valid = int(s.n_valid)
        return o

    def update(self, z) -> None:
        """Merge another StatsRealImagFloat32, or the statistics of a
        complex64 array (computed on the GPU)."""
        other = z if isinstance(z, StatsRealImagFloat32) else _array_stats(z)
        self.real._merge(other.real)
        self.imag._merge(other.imag)
        self.n_valid = self.real.n_valid

    def __str__(self) -> str:
        return f"StatsRealImagFloat32(real={self.real}, imag={self.imag})"

    __repr__ = __str__


# ---------------------------------------------------------------- GPU finisher

def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise errors.HipError("RSLC block finishing needs a ROCm GPU (none visible)")
    return torch


class BlockFinisher:
    """One HIP pass per focused block (include/isce3_amd_rslc.h): scale by
    the range correction, add the block already written, statistics, encode.
    Blocks of up to ``max_pixels`` samples."""

    def __init__(self, max_pixels: int, device: int = 0):
        self._h = None
        h = C.c_void_p()
        _native.check(_native.lib().bp_rslc_finisher_create(int(device), int(max_pixels),
                                                            C.byref(h)))
        self._h = h
        self.device = int(device)
        self.max_pixels = int(max_pixels)
        self._lock = threading.Lock()

    def finish(self, z, range_cor=None, prior=None, data_type: str = "complex64",
               mantissa_nbits: int = 10, stream=None):
        """Finish block ``z`` (2-D complex64, numpy or ROCm torch tensor,
        updated in place like BackgroundWriter.write's ``z``).  Returns
        ``(encoded, stats)``: ``encoded`` is ``z`` itself for the complex64
        types and a new complex32 array (numpy structured ``complex32``, or
        ``torch.complex32`` for a tensor) otherwise."""
        if self._h is None:
            raise errors.RuntimeError_("finisher is closed")
        if data_type not in DATA_TYPES:
            raise ValueError(f"Requested invalid data_type {data_type}")
        if z.ndim != 2:
            raise errors.InvalidArgument("require a 2-D block")
        rows, cols = (int(s) for s in z.shape)
        torch = _torch()
        dev = f"cuda:{self.device}"

        def to_dev(a, shape):
            if a is None:
                return None
            if _is_torch(a):
                t = a.to(device=dev, dtype=torch.complex64).contiguous()
            else:
                t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(dev)
            if tuple(t.shape) != shape:
                raise errors.LengthError(f"expected shape {shape}, got {tuple(t.shape)}")
            return t

        with torch.cuda.device(self.device):
            in_place = (_is_torch(z) and z.is_cuda and z.dtype == torch.complex64 and
                        z.is_contiguous())
            if not _is_torch(z) and (z.dtype != np.complex64 or not z.flags.c_contiguous):
                raise errors.InvalidArgument("z must be a C-contiguous complex64 array")
            if not _is_torch(z) and not z.flags.writeable and (
                    range_cor is not None or prior is not None or
                    data_type == "complex64_zero_mantissa"):
                raise errors.InvalidArgument("z must be a writeable C-contiguous complex64 array")
            dz = z if in_place else to_dev(z, (rows, cols))
            drc = to_dev(range_cor, (cols,))
            dprior = to_dev(prior, (rows, cols))
            denc = None
            if data_type == "complex32":
                denc = torch.empty((rows, cols), dtype=torch.int32, device=dev)
            if stream is None:
                hstream = torch.cuda.current_stream().cuda_stream
            else:
                hstream = stream if isinstance(stream, int) else stream.cuda_stream
            s = _native.StatsRealImagF32_t()
            ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
            with self._lock:
                _native.check(_native.lib().bp_rslc_finish_block(
                    self._h, ptr(dz), rows, cols, ptr(drc), ptr(dprior),
                    DATA_TYPES[data_type], int(mantissa_nbits), ptr(denc), C.byref(s),
                    C.c_void_p(hstream)))
            stats = StatsRealImagFloat32._from_struct(s)
            changed = (range_cor is not None or prior is not None or
                       data_type == "complex64_zero_mantissa")
            if not in_place and changed:
                res = dz.cpu()
                if _is_torch(z):
                    z.copy_(res)
                else:
                    z[...] = res.numpy()
            if denc is None:
                return z, stats
            if _is_torch(z):
                return denc.view(torch.complex32), stats
            return denc.cpu().numpy().view(complex32), stats

    def close(self):
        if self._h is not None:
            _native.lib().bp_rslc_finisher_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


_shared = {}


def _finisher(n: int, device: int = 0) -> BlockFinisher:
    f = _shared.get(device)
    if f is None or f.max_pixels < n:
        f = BlockFinisher(max(n, 1 << 20), device)
        _shared[device] = f
    return f


def _array_stats(z) -> StatsRealImagFloat32:
    if _is_torch(z):
        a = z.reshape(1, -1) if z.ndim != 2 else z
    else:
        a = np.ascontiguousarray(np.asarray(z, dtype=np.complex64)).reshape(1, -1)
    _, stats = _finisher(int(np.prod(a.shape))).finish(a, data_type="complex64")
    return stats


def to_complex32(z):
    """complex64 -> pairs of IEEE half, round to nearest even
    (isce3.core.types.to_complex32, types.py:38-42)."""
    a = z if _is_torch(z) else np.array(z, dtype=np.complex64, order="C", copy=True)
    shape = tuple(a.shape)
    a2 = a.reshape(1, -1) if len(shape) != 2 else a
    if _is_torch(a2):
        a2 = a2.clone()
    enc, _ = _finisher(int(np.prod(shape))).finish(a2, data_type="complex32")
    return enc.reshape(shape)


def truncate_mantissa(z, significant_bits: int = 10) -> None:
    """Zero the low mantissa bits of a complex64 array in place, keeping
    `significant_bits` (isce3.core.types.truncate_mantissa, types.py:116-171)."""
    if not (isinstance(z, np.ndarray) or _is_torch(z)):
        raise ValueError("argument z is not of type np.ndarray")
    if z.dtype not in (np.complex64,) and not (_is_torch(z) and str(z.dtype) == "torch.complex64"):
        raise ValueError("argument z is not complex64 (the GPU encoder's storage type)")
    if significant_bits == 23:
        return
    if not 0 < significant_bits <= 23:
        raise ValueError(f"Require 0 < {significant_bits=} <= 23")
    shape = tuple(z.shape)
    if _is_torch(z):
        view = z.reshape(1, -1)
        _finisher(int(np.prod(shape))).finish(view, data_type="complex64_zero_mantissa",
                                              mantissa_nbits=significant_bits)
        if view.data_ptr() != z.data_ptr():
            z.copy_(view.reshape(shape))
        return
    work = np.ascontiguousarray(z).reshape(1, -1).copy()
    _finisher(work.size).finish(work, data_type="complex64_zero_mantissa",
                                mantissa_nbits=significant_bits)
    z[...] = work.reshape(shape)


# ---------------------------------------------------------------- writer

def unpack_slices(slices):
    """Hashable form of a (rows, cols) pair of slices: one (start, stop, step)
    triple per axis, usable as a dictionary key."""
    return tuple((s.start, s.stop, s.step) for s in slices)


def _decode_c32(a) -> np.ndarray:
    a = np.asarray(a)
    if a.dtype == complex32:
        out = np.empty(a.shape, np.complex64)
        out.real = a["r"].astype(np.float32)
        out.imag = a["i"].astype(np.float32)
        return out
    return a.astype(np.complex64)


# storage type of the dataset each encoding writes to
_STORAGE = {"complex64": np.dtype(np.complex64), "complex64_zero_mantissa": np.dtype(np.complex64),
            "complex32": complex32}


def _stored_dtype(dset):
    """The dataset's element type, or None when the dataset cannot describe
    it (h5py raises TypeError for some compound types, e.g. float16 pairs)."""
    try:
        return np.dtype(dset.dtype)
    except TypeError:
        return None


class BackgroundWriter:
    """Finish and store focused RSLC blocks, optionally in a background thread.

    Same constructor, methods and errors as nisar.workflows.focus
    .BackgroundWriter (focus.py:817-914).  ``write(z, block)`` runs one HIP
    pass over the block (include/isce3_amd_rslc.h): ``z`` scaled in place by
    ``range_cor[cols]``, plus what is already stored at ``block`` when the
    block was written before (mixed-mode focusing visits blocks twice), the
    block's statistics, and its storage encoding; the encoded block goes to
    ``dset.write_direct(..., dest_sel=block)``.  ``queue_write`` runs the same
    call on one worker thread (at most ``num_work_queue`` calls waiting when
    positive); ``notify_finished`` waits for them and re-raises the first
    error one of them raised.  ``stats`` merges the statistics of every block
    written.  Writes are serialised by a lock, so calls from the caller's
    thread and the worker may mix."""

    def __init__(self, range_cor, dset, data_type, mantissa_nbits=10, device: int = 0,
                 num_work_queue: int = 0, timeout: float = 1.0):
        if data_type not in DATA_TYPES:
            raise ValueError(f"Requested invalid data_type {data_type}")
        stored = _stored_dtype(dset)
        fits = stored == _STORAGE[data_type] if stored is not None else data_type == "complex32"
        if not fits:
            raise TypeError(f"Requested {data_type} encoding but the dataset stores "
                            f"{stored if stored is not None else 'an undescribed type'}")
        mantissa_nbits = int(mantissa_nbits)
        if data_type == "complex64_zero_mantissa" and not 0 < mantissa_nbits <= 23:
            raise ValueError(f"Require 0 < significant_bits={mantissa_nbits} <= 23")
        log.debug("RSLC writer: %s encoding (%d mantissa bits), dataset %s", data_type,
                  mantissa_nbits, stored)
        self.range_cor = range_cor
        self.dset = dset
        self.data_type = data_type
        self.mantissa_nbits = mantissa_nbits
        self.device = int(device)
        self._visited_blocks = {}
        self._finisher = None
        self._lock = threading.Lock()
        self._pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="isce3_amd-rslc")
        self._slots = threading.BoundedSemaphore(num_work_queue) if num_work_queue > 0 else None
        self._pending = []
        self._closed = False

    @property
    def stats(self) -> StatsRealImagFloat32:
        total = StatsRealImagFloat32()
        for s in self._visited_blocks.values():
            total.update(s)
        return total

    def block_stats(self, block) -> StatsRealImagFloat32:
        """Statistics of the last write of ``block``."""
        return self._visited_blocks[unpack_slices(block)]

    def readback(self, block) -> np.ndarray:
        """The block as stored, decoded to complex64."""
        return _decode_c32(self.dset[block])

    def _finisher_for(self, pixels: int) -> BlockFinisher:
        if self._finisher is None or self._finisher.max_pixels < pixels:
            self._finisher = BlockFinisher(max(pixels, 1 << 20), self.device)
        return self._finisher

    def write(self, z, block):
        key = unpack_slices(block)
        cols = block[1]
        with self._lock:
            prior = self.readback(block) if key in self._visited_blocks else None
            enc, st = self._finisher_for(int(np.prod(z.shape))).finish(
                z, range_cor=self.range_cor[cols], prior=prior, data_type=self.data_type,
                mantissa_nbits=self.mantissa_nbits)
            self._visited_blocks[key] = st
            if _is_torch(enc):
                import torch
                enc = (enc.view(torch.int32).cpu().numpy().view(complex32)
                       if self.data_type == "complex32" else enc.cpu().numpy())
            self.dset.write_direct(enc, dest_sel=block)
        log.debug("RSLC block %s written%s; real [%g, %g], imag [%g, %g]", key,
                  " (accumulated)" if prior is not None else "", st.real.min, st.real.max,
                  st.imag.min, st.imag.max)

    def _run(self, z, block):
        try:
            self.write(z, block)
        finally:
            if self._slots is not None:
                self._slots.release()

    def queue_write(self, z, block):
        if self._closed:
            raise RuntimeError("queue_write called after notify_finished")
        if self._slots is not None:
            self._slots.acquire()
        self._pending.append(self._pool.submit(self._run, z, block))

    def notify_finished(self):
        """Wait for every queued write; re-raise the first error one raised."""
        self._closed = True
        pending, self._pending = self._pending, []
        errors_ = [f.exception() for f in pending]
        self._pool.shutdown(wait=True)
        first = next((e for e in errors_ if e is not None), None)
        if first is not None:
            raise first

    def __del__(self):
        try:
            self._pool.shutdown(wait=False)
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
