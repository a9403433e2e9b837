"""The oracle on whole seeded scenes: point-target focusing quality (the
acceptance test of tests/python/extensions/pybind/focus/backproject.py:126-170,
re-created on a synthetic scene), the small-scene golden fixture, agreement of
the reference and all-FP64 modes, and the reference's per-pixel failure and
exception rules (Backproject.cpp:88-92,141-171)."""
import os

import numpy as np
import pytest

import oracle
from helpers import analyze_point_target, complex_rel_error, oracle_run, phase_errors, rms, scene_and_raw

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
C = 299792458.0


def point_target_checks(img, sc):
    """Thresholds of backproject.py:160-170 with B = 0.8 fs and the scene's ds."""
    g = sc.out_grid
    t_tgt, r_tgt = sc.target_radar[0]
    row = (t_tgt - g.sensing_start) * g.prf
    col = (r_tgt - g.starting_range) / g.range_pixel_spacing
    r = g.starting_range + g.range_pixel_spacing * np.arange(g.width)
    z = img * np.exp(-1j * 4 * np.pi / g.wavelength * r)[None, :]
    info = analyze_point_target(z, row, col)
    dr, dt = g.range_pixel_spacing, 1.0 / g.prf
    bw_hz = 0.8 * C / (2 * dr)
    range_res = C / (2 * bw_hz)
    # the reference converts azimuth samples to metres with the platform speed
    _, vel = sc.orbit.interpolate(t_tgt)
    az_err = abs(dt * info["azimuth_offset"] * np.linalg.norm(vel))
    az_width = dt * info["azimuth_width"] * np.linalg.norm(vel)
    rg_err = abs(dr * info["range_offset"])
    rg_width = dr * info["range_width"]
    return info, rg_err, az_err, rg_width, az_width, range_res


def test_point_target_quality_oracle():
    sc, raw = scene_and_raw("pt")
    out, h, _, rc = oracle_run("pt", oracle.MODE_REFERENCE)
    assert rc == 0
    np.testing.assert_allclose(h, sc.dem.ref_height, atol=1e-6)
    info, rg_err, az_err, rg_w, az_w, range_res = point_target_checks(out, sc)
    print(info)
    assert rg_err < range_res / 128
    assert az_err < sc.ds / 128
    assert rg_w <= range_res
    assert az_w <= 6.62 / 6.0 * sc.ds


def test_reference_and_fp64_modes_agree():
    for cfg in ("tiny", "small"):
        a, _, _, _ = oracle_run(cfg, oracle.MODE_REFERENCE)
        b, _, _, _ = oracle_run(cfg, oracle.MODE_FP64)
        assert complex_rel_error(a, b) < 1e-6
        assert rms(phase_errors(a, b)) < 1e-6


def test_small_scene_golden_fixture():
    """Config 0 (512 x 2048 pulses -> 256 x 256, 3 targets, flat DEM, noise 0.01)
    against the committed oracle output (every 4th pixel)."""
    fx = np.load(os.path.join(GOLDEN, "small_oracle.npz"))
    out, h, _, rc = oracle_run("small", oracle.MODE_REFERENCE)
    assert rc == 0
    idx = fx["index"]
    np.testing.assert_allclose(out.ravel()[idx], fx["out"], rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(h.ravel()[idx], fx["height"], atol=1e-4)
    assert int(fx["updates"]) == 256 * 256 * 512


def test_failure_rules():
    """Non-bracketing look limits -> rdr2geo fails: out NaN, height NaN, nonzero
    return.  Time limits excluding the solution -> geo2rdr fails: out NaN,
    height finite."""
    sc, raw = scene_and_raw("tiny")
    kw = dict(mode=oracle.MODE_REFERENCE, want_height=True)
    out, h, _, rc = oracle.backproject(sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc,
                                       sc.ds, sc.kernel, "nodelay",
                                       rdr2geo_params={"look_min": 0.0, "look_max": 0.1}, **kw)
    assert rc == 6 and np.isnan(out).all() and np.isnan(h).all()
    t_mid = sc.in_grid.sensing_mid
    out, h, _, rc = oracle.backproject(sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc,
                                       sc.ds, sc.kernel, "nodelay",
                                       geo2rdr_params={"time_start": t_mid + 20.0,
                                                       "time_end": t_mid + 30.0}, **kw)
    assert rc == 6 and np.isnan(out).all() and np.isfinite(h).all()


def test_output_outside_orbit_raises():
    sc, raw = scene_and_raw("tiny")
    g = sc.out_grid.copy()
    g.sensing_start = sc.orbit.end_time + 5.0
    with pytest.raises(oracle.OracleError):
        oracle.backproject(sc.out_geometry(g), raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                           sc.kernel, "nodelay")


def test_tsx_delay_changes_phase_consistently():
    """With the TSX model the focused phase shifts by 2 pi fc tau_atm (plus the
    resampling of tau_atm / dtau bins); magnitude stays focused."""
    sc, raw = scene_and_raw("tiny")
    a, _, _, _ = oracle.backproject(sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc,
                                    sc.ds, sc.kernel, "nodelay")
    b, _, diag, _ = oracle.backproject(sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc,
                                       sc.ds, sc.kernel, "tsx", want_diag=True)
    tau = np.array([d.tau_atm for d in diag]).reshape(a.shape)
    assert np.all(tau > 2 * 2.3 * np.exp(-300 / 6000) / C) and np.all(tau < 2 * 2.3 * 1.2 / C)
