"""CPU oracle of range compression -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's CPU-baseline leg use this module, as the checker
of the HIP range compressor.  It restates, in numpy, the reference's
frequency-domain matched filter:

  next_fast_power     cxx/isce3/fft/FFTUtil.icc:37-73
  form_linear_chirp   cxx/isce3/focus/Chirp.cpp:10-52
  RangeComp           cxx/isce3/focus/RangeComp.cpp:10-131 (output sizes :10-20,
                      matched filter :22-37, FFT size :59, first valid sample
                      :81-91, compression :93-131)

`precision="fp64"` runs the same steps in complex128 (the yardstick the GPU
result's error is stated against); `"fp32"` keeps the reference's complex64
buffers (numpy's pocketfft in single precision stands in for FFTW).
"""
from __future__ import annotations

import math

import numpy as np

FULL, VALID, SAME = 0, 1, 2
_MODES = {"full": FULL, "valid": VALID, "same": SAME}


def next_fast_power(n: int) -> int:
    """Smallest m = 2^a 3^b 5^c >= n (FFTUtil.icc:37-73)."""
    if n < 0:
        raise ValueError("input must be non-negative")
    if n <= 1:
        return 1
    logn = math.log(n)
    max5 = math.ceil(logn / math.log(5))
    max3 = math.ceil(logn / math.log(3))
    best = None
    for x5 in range(max5 + 1):
        for x3 in range(max3 + 1):
            m = 5 ** x5 * 3 ** x3
            while m < n:
                m *= 2
            if best is None or m < best:
                best = m
    return best


def form_linear_chirp(chirprate: float, duration: float, samplerate: float,
                      centerfreq: float = 0.0, amplitude: float = 1.0,
                      phi: float = 0.0) -> np.ndarray:
    """Time-domain LFM chirp, complex64, odd sample count (Chirp.cpp:10-52)."""
    if duration <= 0:
        raise ValueError("chirp duration must be > 0")
    if samplerate <= 0:
        raise ValueError("sampling rate must be > 0")
    if amplitude <= 0:
        raise ValueError("amplitude must be > 0")
    size = int(math.floor(samplerate * duration + 1))
    if size % 2 == 0:
        size += 1
    spacing = 1.0 / samplerate
    tau = -0.5 * (size - 1) * spacing + spacing * np.arange(size)
    phase = phi + 2.0 * np.pi * (centerfreq + 0.5 * chirprate * tau) * tau
    return (amplitude * np.exp(1j * phase)).astype(np.complex64)


def output_size(m: int, n: int, mode: int) -> int:
    """RangeComp.cpp:10-20 (m = chirp size, n = input size)."""
    if mode == FULL:
        return m + n - 1
    if mode == VALID:
        return max(m, n) - min(m, n) + 1
    return n


def first_valid_sample(m: int, mode: int) -> int:
    """RangeComp.cpp:81-91."""
    if mode == FULL:
        return m - 1
    if mode == VALID:
        return 0
    return m // 2


def rangecompress(chirp: np.ndarray, x: np.ndarray, mode: int = FULL,
                  precision: str = "fp64") -> np.ndarray:
    """Compress each row of `x` (batch x n) with the matched filter of
    `chirp`: FFT of the zero-padded input times the FFT of the time-reversed
    conjugate chirp, scaled by 1/fft_size, inverse FFT, cropped to the mode's
    output window (RangeComp.cpp:22-37, 93-131)."""
    x = np.atleast_2d(np.asarray(x))
    m, n = len(chirp), x.shape[1]
    nfft = next_fast_power(output_size(m, n, FULL))
    ct = np.complex128 if precision == "fp64" else np.complex64
    ref = np.zeros(nfft, ct)
    ref[:m] = np.conj(np.asarray(chirp, ct)[::-1])
    ref = np.fft.fft(ref)
    buf = np.zeros((x.shape[0], nfft), ct)
    buf[:, :n] = x
    spec = np.fft.fft(buf, axis=1)
    spec = (spec * (ref * (1.0 / nfft)).astype(ct)).astype(ct)
    y = np.fft.ifft(spec, axis=1) * nfft      # numpy's ifft divides by nfft
    off = 0 if mode == FULL else (m - 1 if mode == VALID else m // 2)
    out = y[:, off:off + output_size(m, n, mode)]
    return out.astype(np.complex64 if precision == "fp32" else np.complex128)


def direct_correlation(chirp: np.ndarray, x: np.ndarray, mode: int = FULL) -> np.ndarray:
    """The same output by direct summation in complex128: the full linear
    convolution of each row with conj(chirp[::-1]), cropped like
    rangecompress.  Independent of any FFT: it pins the matched filter, the
    padding and the crop offsets."""
    x = np.atleast_2d(np.asarray(x, np.complex128))
    h = np.conj(np.asarray(chirp, np.complex128)[::-1])
    m, n = len(h), x.shape[1]
    full = np.stack([np.convolve(r, h) for r in x])
    off = 0 if mode == FULL else (m - 1 if mode == VALID else m // 2)
    return full[:, off:off + output_size(m, n, mode)]
