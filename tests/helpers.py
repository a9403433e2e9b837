"""Shared helpers of the test suite: error metrics and cached scenes."""
from __future__ import annotations

import functools

import numpy as np


def complex_rel_error(z: np.ndarray, ref: np.ndarray) -> float:
    m = np.isfinite(ref) & np.isfinite(z)
    return float(np.linalg.norm(z[m] - ref[m]) / np.linalg.norm(ref[m]))


def phase_errors(z: np.ndarray, ref: np.ndarray, rel_floor: float = 1e-2):
    """arg(z * conj(ref)) over pixels with |ref| >= rel_floor * max|ref| (rad)."""
    a = np.abs(ref)
    m = np.isfinite(a) & np.isfinite(z) & (a >= rel_floor * np.nanmax(a))
    return np.angle(z[m] * np.conj(ref[m]))


def rms(x) -> float:
    x = np.asarray(x)
    return float(np.sqrt(np.mean(x ** 2))) if x.size else 0.0


@functools.lru_cache(maxsize=None)
def scene_and_raw(config: str, **kw):
    from isce3_amd import sim
    sc = sim.make_scene(config, **kw)
    raw = sim.simulate(sc)
    raw.setflags(write=False)
    return sc, raw


@functools.lru_cache(maxsize=None)
def oracle_run(config: str, mode: int, want_diag: bool = False):
    import oracle
    sc, raw = scene_and_raw(config)
    out, h, diag, rc = oracle.backproject(sc.out_geometry(), raw, sc.in_geometry(), sc.dem,
                                          sc.fc, sc.ds, sc.kernel, sc.dry_tropo_model,
                                          mode=mode, kernel_bandwidth=0.8, want_height=True,
                                          want_diag=want_diag)
    return out, h, diag, rc


def _upsample2(z: np.ndarray, nov: int) -> np.ndarray:
    """Band-limited 2-D upsampling by zero padding the centred spectrum."""
    n0, n1 = z.shape
    Z = np.fft.fftshift(np.fft.fft2(z))
    P = np.zeros((n0 * nov, n1 * nov), dtype=complex)
    o0, o1 = (n0 * nov - n0) // 2, (n1 * nov - n1) // 2
    P[o0:o0 + n0, o1:o1 + n1] = Z
    return np.fft.ifft2(np.fft.ifftshift(P)) * nov * nov


def _width_3db(cut: np.ndarray, ipk: int) -> float:
    p = np.abs(cut) ** 2
    half = 0.5 * p[ipk]
    i = ipk
    while i > 0 and p[i] >= half:
        i -= 1
    left = i + (half - p[i]) / (p[i + 1] - p[i])
    j = ipk
    while j < p.size - 1 and p[j] >= half:
        j += 1
    right = j - 1 + (p[j - 1] - half) / (p[j - 1] - p[j])
    return right - left


def analyze_point_target(img: np.ndarray, row: float, col: float, half: int = 16,
                         nov: int = 32):
    """Locate the response nearest (row, col) to 1/nov pixel and measure its
    -3 dB widths in range (columns) and azimuth (rows), in pixels.  Returns
    dict(range_offset, azimuth_offset, range_width, azimuth_width, peak)."""
    r0, c0 = int(round(row)), int(round(col))
    chip = img[r0 - half:r0 + half, c0 - half:c0 + half]
    up = _upsample2(chip, nov)
    a = np.abs(up)
    i, j = np.unravel_index(np.argmax(a), a.shape)
    pk_row = r0 - half + i / nov
    pk_col = c0 - half + j / nov
    return {"azimuth_offset": pk_row - row, "range_offset": pk_col - col,
            "range_width": _width_3db(up[i, :], j) / nov,
            "azimuth_width": _width_3db(up[:, j], i) / nov, "peak": float(a[i, j] / nov ** 0)}
