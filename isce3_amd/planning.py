"""Block planning of the focusing workflow: the caller's tiling of the output
grid and the raw-data time span each tile needs.

Look-alikes of ``isce3.focus.get_radar_velocities`` / ``get_sar_duration``
(python/packages/isce3/focus/sar_duration.py:9-69) and of
``nisar.workflows.focus.plan_processing_blocks`` / ``total_bounds`` /
``is_overlapping`` / ``get_rdr2geo_params`` / ``get_geo2rdr_params``
(python/packages/nisar/workflows/focus.py:665-795).  ``cfg`` is any object
with the runconfig attributes these read (``cfg.processing.azcomp.block_size
.{azimuth,range}``, ``cfg.processing.azcomp.azimuth_resolution``,
``cfg.processing.rdr2geo.{tol_height,look_min_deg,look_max_deg}``,
``cfg.processing.geo2rdr.{tol_aztime,time_start,time_end}``), e.g. nested
``types.SimpleNamespace`` objects.
"""
from __future__ import annotations

import logging
from typing import Optional

import numpy as np

from . import core, geometry

log = logging.getLogger("isce3_amd.planning")


def get_radar_velocities(orbit, time: Optional[float] = None,
                         ellipsoid: Optional[core.Ellipsoid] = None):
    """(vs, vg): platform along-track speed and ground-track speed at `time`
    (orbit midpoint when None) (sar_duration.py:9-44).  As in the reference,
    ``vs = |vel - vel.dot(nadir)|`` subtracts the scalar projection from every
    component."""
    ellipsoid = ellipsoid if ellipsoid is not None else core.Ellipsoid()
    t = time if time is not None else orbit.mid_time
    pos, vel = orbit.interpolate(t)
    tcn = core.geodetic_tcn(pos, vel, ellipsoid)
    nadir = tcn.x2
    vs = float(np.linalg.norm(vel - vel.dot(nadir)))
    lon, lat, h = ellipsoid.xyz_to_lon_lat(pos)
    hdg = geometry.heading(lon, lat, vel)
    a = ellipsoid.r_dir(hdg, lat)
    return vs, vs / (1 + h / a)


def get_sar_duration(t: float, r: float, orbit, ellipsoid, azres: float,
                     wavelength: float) -> float:
    """Approximate synthetic-aperture duration (s) for azimuth resolution
    `azres` at (t, r): wavelength / (2 azres) * r / vg (sar_duration.py:46-69)."""
    _, vg = get_radar_velocities(orbit, t, ellipsoid)
    return wavelength / (2 * azres) * r / vg


def get_rdr2geo_params(cfg) -> dict:
    """focus.py:665-685: tol_height, look_min/look_max in radians."""
    g = cfg.processing.rdr2geo
    return dict(tol_height=g.tol_height, look_min=np.deg2rad(g.look_min_deg),
                look_max=np.deg2rad(g.look_max_deg))


def get_geo2rdr_params(cfg, orbit=None) -> dict:
    """focus.py:688-719: geo2rdr parameters, time bounds defaulting to the
    orbit's span when an orbit is given."""
    geo2rdr_params = dict(vars(cfg.processing.geo2rdr))
    if orbit is not None:
        if geo2rdr_params.get("time_start", None) is None:
            geo2rdr_params["time_start"] = orbit.start_time
        if geo2rdr_params.get("time_end", None) is None:
            geo2rdr_params["time_end"] = orbit.end_time
    return geo2rdr_params


def plan_processing_blocks(cfg, grid, doppler, dem, orbit, pad: float = 0.1):
    """Tile the output grid into blocks of ``cfg.processing.azcomp.block_size``
    (range < 1: the whole width) and find the raw-data time bounds of each:
    the zero-to-native-Doppler reskew time of the four corners, widened by half
    the padded synthetic-aperture duration at the block's mid time and far
    range (focus.py:726-783).  Returns ``[((rows, cols), (t0, t1)), ...]``."""
    assert pad >= 0.0
    ac = cfg.processing.azcomp
    nr = ac.block_size.range
    na = ac.block_size.azimuth
    if nr < 1:
        nr = grid.width
    blocks = []
    for i in range(0, grid.length, na):
        imax = min(i + na, grid.length)
        for j in range(0, grid.width, nr):
            jmax = min(j + nr, grid.width)
            blocks.append(np.s_[i:imax, j:jmax])
    zerodop = core.LUT2d()
    r2g = get_rdr2geo_params(cfg)
    g2r = get_geo2rdr_params(cfg, orbit)
    results = []
    for rows, cols in blocks:
        raw_times = []
        for (u, v) in ((rows.start, cols.start), (rows.stop - 1, cols.start),
                       (rows.start, cols.stop - 1), (rows.stop - 1, cols.stop - 1)):
            t = grid.sensing_start + u / grid.prf
            r = grid.starting_range + v * grid.range_pixel_spacing
            try:
                traw, _ = geometry.rdr2rdr(t, r, orbit, grid.lookside, zerodop,
                                           grid.wavelength, dem, doppler_out=doppler,
                                           rdr2geo_params=r2g, geo2rdr_params=g2r)
            except RuntimeError as e:
                log.error(f"Reskew zero-to-native failed at t={t} r={r}")
                raise RuntimeError("Could not compute imaging geometry") from e
            raw_times.append(traw)
        sub_grid = grid[rows, cols]
        cpi = get_sar_duration(sub_grid.sensing_mid, sub_grid.end_range, orbit,
                               core.Ellipsoid(), ac.azimuth_resolution, sub_grid.wavelength)
        cpi *= 1.0 + pad
        results.append(((rows, cols), (min(raw_times) - cpi / 2, max(raw_times) + cpi / 2)))
    return results


def total_bounds(blocks_bounds):
    """Union of the blocks' time bounds (focus.py:786-789)."""
    return (min(t0 for _, (t0, t1) in blocks_bounds),
            max(t1 for _, (t0, t1) in blocks_bounds))


def is_overlapping(a, b, c, d) -> bool:
    """[a, b] and [c, d] intersect (focus.py:792-794)."""
    assert (b >= a) and (d >= c)
    return (d >= a) and (c <= b)
