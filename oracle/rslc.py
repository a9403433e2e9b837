"""Oracle (test infrastructure only): CPU restatement of the per-block work of
the RSLC writer.  Imported by tests/ and bench.py's CPU baseline, never by
the isce3_amd package.

Follows
  * nisar/workflows/focus.py:883-914 (BackgroundWriter.write): z *= range_cor,
    z += readback when the block was written before, statistics, encode;
  * isce3/math/Stats.cpp:33-62 (Stats<T>::update(value), Welford),
    :64-96 (update(other), Chan), :116-124 (sample_stddev), :128-137
    (StatsRealImag skips a sample when either part is NaN);
  * isce3/core/types.py:38-42 (to_complex32: float32 -> float16 casts) and
    :116-171 (truncate_mantissa: AND with a mask of the high bits).

Two statistics restatements: ``stats_welford`` replays the reference's
sequential update sample by sample (float32 running mean, FP64 provisional
mean and squared-difference sum) for small inputs, ``stats_exact`` gives the
FP64 two-pass values any order of summation approaches, for large inputs.
The complex product is rounded term by term, (ac - bd) + i(ad + bc) with
each product and sum rounded to float32 (no fused multiply-add).
"""
from __future__ import annotations

import math

import numpy as np

complex32 = np.dtype([("r", np.float16), ("i", np.float16)])
MANT_BITS = 23


class Stats:
    """State of Stats<float> (Stats.h:15-49)."""

    def __init__(self):
        self.min = math.nan
        self.max = math.nan
        self.mean = np.float32(0.0)
        self.n_valid = 0
        self.real_valued_mean = 0.0
        self.square_diff_sum = 0.0

    def update_value(self, value: np.float32) -> None:
        # Stats.cpp:33-62
        if math.isnan(value):
            return
        self.n_valid += 1
        value = np.float32(value)
        self.mean = np.float32(self.mean + np.float32(
            np.float32(value - self.mean) / np.float32(self.n_valid)))
        delta = float(value) - self.real_valued_mean
        self.real_valued_mean += delta / self.n_valid
        self.square_diff_sum += delta * (float(value) - self.real_valued_mean)
        if math.isnan(self.max) or value > self.max:
            self.max = float(value)
        if math.isnan(self.min) or value < self.min:
            self.min = float(value)

    def update(self, other: "Stats") -> None:
        # Stats.cpp:64-96
        if other.n_valid <= 0:
            return
        nself, ntotal = self.n_valid, self.n_valid + other.n_valid
        self.n_valid += other.n_valid
        self.mean = np.float32(self.mean + np.float32(
            np.float32(other.mean - self.mean) * np.float32(other.n_valid / np.float32(ntotal))))
        delta = other.real_valued_mean - self.real_valued_mean
        self.real_valued_mean += delta * other.n_valid / ntotal
        self.square_diff_sum += other.square_diff_sum + delta ** 2 * nself * other.n_valid / ntotal
        if math.isnan(self.max) or other.max > self.max:
            self.max = other.max
        if math.isnan(self.min) or other.min < self.min:
            self.min = other.min

    @property
    def sample_stddev(self) -> float:
        # Stats.cpp:116-124
        if self.n_valid > 1:
            return math.sqrt(self.square_diff_sum / (self.n_valid - 1))
        return 0.0


def stats_welford(z) -> tuple[Stats, Stats]:
    """StatsRealImag<float>(values) sample by sample (Stats.cpp:128-157)."""
    z = np.asarray(z, dtype=np.complex64).ravel()
    re, im = Stats(), Stats()
    for v in z:
        if math.isnan(v.real) or math.isnan(v.imag):
            continue
        re.update_value(np.float32(v.real))
        im.update_value(np.float32(v.imag))
    return re, im


def stats_exact(z) -> dict:
    """FP64 two-pass statistics of the valid samples of complex array z:
    {n_valid, real: (min, max, mean, square_diff_sum), imag: (...)}."""
    z = np.asarray(z, dtype=np.complex64).ravel()
    ok = ~(np.isnan(z.real) | np.isnan(z.imag))
    out = {"n_valid": int(ok.sum())}
    for name, part in (("real", z.real[ok]), ("imag", z.imag[ok])):
        if part.size == 0:
            out[name] = (math.nan, math.nan, 0.0, 0.0)
            continue
        x = part.astype(np.float64)
        m = float(x.mean())
        out[name] = (float(part.min()), float(part.max()), m, float(((x - m) ** 2).sum()))
    return out


def cmul_rounded(z, c) -> np.ndarray:
    """complex64 product with each real product and sum rounded to float32."""
    z = np.asarray(z, dtype=np.complex64)
    c = np.asarray(c, dtype=np.complex64)
    a, b = z.real.astype(np.float32), z.imag.astype(np.float32)
    cr, ci = c.real.astype(np.float32), c.imag.astype(np.float32)
    out = np.empty(np.broadcast(z, c).shape, np.complex64)
    out.real = np.float32(a * cr) - np.float32(b * ci)
    out.imag = np.float32(a * ci) + np.float32(b * cr)
    return out


def to_complex32(z) -> np.ndarray:
    """types.py:38-42: each part cast to float16 (round to nearest even)."""
    z = np.asarray(z, dtype=np.complex64)
    zf = np.zeros(z.shape, dtype=complex32)
    with np.errstate(over="ignore"):
        zf["r"] = z.real
        zf["i"] = z.imag
    return zf


def truncate_mantissa(z, significant_bits: int = 10) -> np.ndarray:
    """types.py:116-171 on a copy of complex64 z: keep the top
    `significant_bits` of each part's 23-bit mantissa."""
    z = np.array(z, dtype=np.complex64, copy=True)
    if significant_bits == MANT_BITS:
        return z
    if not 0 < significant_bits <= MANT_BITS:
        raise ValueError(f"Require 0 < {significant_bits=} <= {MANT_BITS}")
    mask = np.uint32(((1 << 32) - 1) << (MANT_BITS - significant_bits) & ((1 << 32) - 1))
    u = z.view(np.uint32)
    u &= mask
    return z


def finish_block(z, range_cor=None, prior=None, data_type="complex64", mantissa_nbits=10):
    """BackgroundWriter.write on arrays: returns (z after scaling and
    accumulation, encoded block, exact statistics)."""
    z = np.array(z, dtype=np.complex64, copy=True)
    if range_cor is not None:
        z = cmul_rounded(z, np.asarray(range_cor, np.complex64)[None, :])
    if prior is not None:
        p = np.asarray(prior, dtype=np.complex64)
        zz = np.empty_like(z)
        zz.real = z.real + p.real
        zz.imag = z.imag + p.imag
        z = zz
    stats = stats_exact(z)
    if data_type == "complex64":
        enc = z
    elif data_type == "complex32":
        enc = to_complex32(z)
    elif data_type == "complex64_zero_mantissa":
        z = truncate_mantissa(z, mantissa_nbits)
        enc = z
    else:
        raise ValueError(f"Requested invalid data_type {data_type}")
    return z, enc, stats
