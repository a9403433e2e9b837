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
