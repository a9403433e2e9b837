"""Range compression on the GPU: look-alike of isce3.focus.RangeComp and
isce3.focus.form_linear_chirp.

RangeComp mirrors the binding python/extensions/pybind_isce3/focus/RangeComp.cpp:
37-101 (constructor ``RangeComp(chirp, inputsize, maxbatch=1, mode=Mode.Full)``,
``rangecompress(out, in)`` with the binding's 1-D/2-D shape rules and
``std::length_error``/``std::invalid_argument`` checks, and the read-only
properties).  The compression runs in the fused FFT-convolution HIP kernel
behind include/isce3_amd_rangecomp.h; ``out`` and ``in`` may be numpy arrays
(staged through device buffers) or ROCm torch tensors (used in place).
"""
from __future__ import annotations

import ctypes as C
import enum
import math

import numpy as np

from . import _native, errors


class Mode(enum.IntEnum):
    """RangeComp::Mode (cxx/isce3/focus/RangeComp.h:13-40)."""
    Full = 0
    Valid = 1
    Same = 2


def form_linear_chirp(chirprate: float, duration: float, samplerate: float,
                      centerfreq: float = 0.0, amplitude: float = 1.0,
                      phi: float = 0.0) -> np.ndarray:
    """Time-domain LFM chirp replica (cxx/isce3/focus/Chirp.cpp:10-52): an odd
    number of samples floor(samplerate * duration + 1) (rounded up to odd),
    centred on t = 0, phase phi + 2 pi (centerfreq + chirprate t / 2) t."""
    if duration <= 0.0:
        raise errors.DomainError("chirp duration must be > 0")
    if samplerate <= 0.0:
        raise errors.DomainError("sampling rate must be > 0")
    if amplitude <= 0.0:
        raise errors.DomainError("amplitude must be > 0")
    d_size = samplerate * duration + 1
    if d_size > 2 ** 31 - 1:
        raise OverflowError("chirp size exceeds max int value")
    size = int(math.floor(d_size))
    if size % 2 == 0:
        size += 1
    spacing = 1.0 / samplerate
    tau = -0.5 * (size - 1) * spacing + spacing * np.arange(size)
    phase = phi + 2.0 * np.pi * (centerfreq + 0.5 * chirprate * tau) * tau
    return (amplitude * np.exp(1j * phase)).astype(np.complex64)


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


class RangeComp:
    """Range compression processor (cxx/isce3/focus/RangeComp.h:11-106)."""

    Mode = Mode

    def __init__(self, chirp, inputsize: int, maxbatch: int = 1, mode: Mode = Mode.Full,
                 device: int = 0):
        self._h = None
        chirp = np.ascontiguousarray(np.asarray(chirp, dtype=np.complex64).ravel())
        if chirp.size > 2 ** 31 - 1:
            raise OverflowError("chirp length exceeds max int")
        h = C.c_void_p()
        _native.check(_native.lib().bp_rangecomp_create(
            int(device), chirp.ctypes.data, int(chirp.size), int(inputsize), int(maxbatch),
            int(mode), C.byref(h)))
        self._h = h
        self._device = int(device)
        sizes = (C.c_int32 * 7)()
        _native.check(_native.lib().bp_rangecomp_sizes(h, sizes))
        (self._chirp_size, self._input_size, self._fft_size, self._maxbatch, m,
         self._output_size, self._first_valid) = list(sizes)
        self._mode = Mode(m)
        plan = (C.c_int32 * 3)()
        _native.check(_native.lib().bp_rangecomp_plan(h, plan))
        self._plan = (("fused", "overlap-save", "four-step")[plan[0]], int(plan[1]), int(plan[2]))

    chirp_size = property(lambda self: self._chirp_size)
    input_size = property(lambda self: self._input_size)
    fft_size = property(lambda self: self._fft_size)
    maxbatch = property(lambda self: self._maxbatch)
    mode = property(lambda self: self._mode)
    output_size = property(lambda self: self._output_size)
    first_valid_sample = property(lambda self: self._first_valid)
    #: how the pulses run: (kind, transform length, transforms per pulse),
    #: kind "fused" (one transform per pulse), "overlap-save" or "four-step"
    plan = property(lambda self: self._plan)

    def rangecompress(self, out, in_, stream=None) -> None:
        """Compress a batch of pulses: ``in_`` of shape (batch, input_size) or
        (input_size,), ``out`` of shape (batch, output_size) or
        (output_size,), complex64, written in place."""
        if self._h is None:
            raise errors.RuntimeError_("range compressor is closed")
        if in_.ndim != out.ndim:
            raise errors.LengthError("require same ndim on input and output")
        if in_.ndim == 2:
            batch = int(in_.shape[0])
            if in_.shape[0] != out.shape[0]:
                raise errors.LengthError("require equal batch size on input and output")
            if in_.shape[1] != self._input_size:
                raise errors.LengthError("unexpected input length")
            if out.shape[1] != self._output_size:
                raise errors.LengthError("unexpected output length")
        elif in_.ndim == 1:
            batch = 1
            if in_.shape[0] != self._input_size:
                raise errors.LengthError("unexpected input length")
            if out.shape[0] != self._output_size:
                raise errors.LengthError("unexpected output length")
        else:
            raise errors.InvalidArgument("require 1D or 2D data")
        if batch > self._maxbatch:
            raise errors.LengthError("batch size exceeds max batch")
        import torch
        if not torch.cuda.is_available():
            raise errors.HipError("range compression needs a ROCm GPU (none visible)")
        dev = f"cuda:{self._device}"
        with torch.cuda.device(self._device):
            if _is_torch(in_):
                d_in = in_.to(dev).contiguous().to(torch.complex64)
            else:
                d_in = torch.from_numpy(np.ascontiguousarray(in_, dtype=np.complex64)).to(dev)
            in_place = (_is_torch(out) and out.is_cuda and out.dtype == torch.complex64
                        and out.is_contiguous())
            if not _is_torch(out):
                if out.dtype != np.complex64 or not out.flags.c_contiguous or \
                        not out.flags.writeable:
                    raise errors.InvalidArgument(
                        "out must be a writeable C-contiguous complex64 array")
            d_out = out if in_place else torch.empty(tuple(out.shape), dtype=torch.complex64,
                                                     device=dev)
            if stream is None:
                hstream = torch.cuda.current_stream().cuda_stream
            else:
                hstream = stream if isinstance(stream, int) else stream.cuda_stream
            _native.check(_native.lib().bp_rangecomp_run(
                self._h, C.c_void_p(d_out.data_ptr()), C.c_void_p(d_in.data_ptr()), batch,
                C.c_void_p(hstream)))
            if not in_place:
                res = d_out.cpu()
                if _is_torch(out):
                    out.copy_(res)
                else:
                    out[...] = res.numpy()

    def close(self):
        if self._h is not None:
            _native.lib().bp_rangecomp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
