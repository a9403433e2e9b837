"""GPU edge cases against the oracle: failure rules, exceptions, TSX delay,
non-zero Doppler LUTs, outputs reaching past the data, resampled output grids,
device-tensor I/O, the small-scene golden fixture and point-target quality."""
import os

import numpy as np
import pytest

import oracle
from helpers import analyze_point_target, complex_rel_error, phase_errors, rms, scene_and_raw

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
C = 299792458.0


def gpu(sc, raw, out_grid=None, precision="plain", in_geom=None, out_geom=None, **kw):
    from isce3_amd.focus import backproject
    g = out_grid or sc.out_grid
    out = np.empty(g.shape, np.complex64)
    h = np.empty(g.shape, np.float32)
    err = backproject(out, out_geom or sc.out_geometry(g), raw, in_geom or sc.in_geometry(),
                      sc.dem, sc.fc, sc.ds, sc.kernel, kw.pop("tropo", sc.dry_tropo_model),
                      height=h, precision=precision, **kw)
    return out, h, err


def orc(sc, raw, out_grid=None, mode=0, in_geom=None, out_geom=None, **kw):
    g = out_grid or sc.out_grid
    out, h, _, rc = oracle.backproject(out_geom or sc.out_geometry(g), raw,
                                       in_geom or sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                                       sc.kernel, kw.pop("tropo", sc.dry_tropo_model), mode=mode,
                                       kernel_bandwidth=0.8, want_height=True, **kw)
    return out, h, rc


def assert_plain_parity(z, ref):
    both_nan = np.isnan(z) == np.isnan(ref)
    assert both_nan.all(), "NaN pattern differs from the oracle"
    if np.isfinite(ref).any() and np.abs(ref[np.isfinite(ref)]).max() > 0:
        assert complex_rel_error(z, ref) <= 1e-5
        assert rms(phase_errors(z, ref)) <= 1e-5


def test_failure_rules_match_oracle(torch_gpu):
    sc, raw = scene_and_raw("tiny")
    r2g = {"look_min": 0.0, "look_max": 0.1}
    z, h, err = gpu(sc, raw, rdr2geo_params=r2g)
    ref, href, rc = orc(sc, raw, rdr2geo_params=r2g)
    assert err and rc == 6 and np.isnan(z).all() and np.isnan(h).all() and np.isnan(href).all()
    t_mid = sc.in_grid.sensing_mid
    g2r = {"time_start": t_mid + 20.0, "time_end": t_mid + 30.0}
    z, h, err = gpu(sc, raw, geo2rdr_params=g2r)
    ref, href, rc = orc(sc, raw, geo2rdr_params=g2r)
    assert err and rc == 6 and np.isnan(z).all() and np.isfinite(h).all()
    np.testing.assert_allclose(h, href, atol=1e-3)


def test_partial_failure_pattern(torch_gpu):
    """A look interval that brackets the root only for the nearer ranges fails
    the far part of the grid; NaN maps agree with the oracle pixel by pixel."""
    sc, raw = scene_and_raw("tiny")
    from isce3_amd.geometry import rdr2geo_bracket
    g = sc.out_grid
    x_far = rdr2geo_bracket(g.sensing_mid, g.slant_range(g.width * 0.6), sc.orbit, "right", 0.0,
                            g.wavelength, sc.dem)
    p, _ = sc.orbit.interpolate(g.sensing_mid)
    look = np.arccos(np.dot(-p / np.linalg.norm(p), (x_far - p) / np.linalg.norm(x_far - p)))
    r2g = {"look_min": 0.2, "look_max": float(look)}
    z, h, err = gpu(sc, raw, rdr2geo_params=r2g)
    ref, href, rc = orc(sc, raw, rdr2geo_params=r2g)
    assert err and rc == 6
    assert 0 < np.isnan(ref).sum() < ref.size
    assert_plain_parity(z, ref)
    np.testing.assert_array_equal(np.isnan(h), np.isnan(href))


def test_output_outside_orbit_raises(torch_gpu):
    sc, raw = scene_and_raw("tiny")
    g = sc.out_grid.copy()
    g.sensing_start = sc.orbit.end_time + 5.0
    with pytest.raises(IndexError):
        gpu(sc, raw, out_grid=g)


def test_pulses_outside_orbit_raise(torch_gpu):
    sc, raw = scene_and_raw("tiny")
    from isce3_amd.container import RadarGeometry
    ig = sc.in_grid.copy()
    ig.sensing_start = sc.orbit.start_time - 1.0
    with pytest.raises(IndexError):
        gpu(sc, raw, in_geom=RadarGeometry(ig, sc.orbit, sc.doppler))


@pytest.mark.parametrize("precision", ["plain", "split"])
def test_tsx_troposphere(torch_gpu, precision):
    sc, raw = scene_and_raw("tiny")
    z, h, err = gpu(sc, raw, precision=precision, tropo="tsx")
    ref, _, rc = orc(sc, raw, mode=0 if precision == "plain" else 1, tropo="tsx")
    assert not err and rc == 0
    tol = 1e-5 if precision == "plain" else 1e-3
    assert complex_rel_error(z, ref) <= tol
    assert rms(phase_errors(z, ref)) <= tol


@pytest.mark.parametrize("squint", ["mild", "strong"])
@pytest.mark.parametrize("precision", ["plain", "split"])
def test_nonzero_doppler_luts(torch_gpu, precision, squint):
    """Doppler LUTs varying in azimuth time and slant range: 11-19 Hz, and
    500-1100 Hz (squint ~0.3-0.5 deg, targets ~10 km ahead of broadside) over a
    raster DEM that reaches there.  The oracle's squinted geometry is pinned by
    the invariants of test_oracle_pins.test_squinted_geometry_invariants."""
    import dataclasses
    from isce3_amd import sim
    from isce3_amd.container import RadarGeometry
    from isce3_amd.core import LUT2d
    from isce3_amd.geometry import DEMInterpolator
    sc, raw = scene_and_raw("tiny")
    g = sc.in_grid
    xs = g.starting_range + np.linspace(-2e3, g.width * g.range_pixel_spacing + 2e3, 6)
    ys = g.sensing_start + np.linspace(-1.0, g.length / g.prf + 1.0, 5)
    if squint == "mild":
        data = 15.0 + 4.0 * np.sin(np.add.outer(ys * 0.7, xs * 1e-4))
    else:
        data = 800.0 + 300.0 * np.sin(np.add.outer((ys - ys[0]) * 0.9, (xs - xs[0]) * 2e-4))
        dem = DEMInterpolator(0.0, method="bilinear")
        step = sc.dem.delta_x
        dem.load_dem_array(sim.fractal_dem(1024, 0.8, 200.0, seed=21),
                           sc.dem.x_start - 200 * step, -0.05, step, step, 4326)
        sc = dataclasses.replace(sc, dem=dem)
    dop = LUT2d(xs, ys, data)
    ig = RadarGeometry(g, sc.orbit, dop)
    og = RadarGeometry(sc.out_grid, sc.orbit, dop)
    z, h, err = gpu(sc, raw, precision=precision, in_geom=ig, out_geom=og)
    ref, href, rc = orc(sc, raw, mode=0 if precision == "plain" else 1, in_geom=ig, out_geom=og)
    assert err == (rc == 6)
    tol = 1e-5 if precision == "plain" else 1e-3
    assert complex_rel_error(z, ref) <= tol
    assert rms(phase_errors(z, ref)) <= tol
    np.testing.assert_allclose(h, href, atol=1e-3)


@pytest.mark.parametrize("precision", ["plain", "split"])
def test_output_grid_past_the_data(torch_gpu, precision):
    """Range samples beyond both ends of the swath (zero-padded taps) and lines
    whose aperture is clipped by, or lies wholly outside, the pulse block."""
    sc, raw = scene_and_raw("tiny")
    from isce3_amd.product import RadarGridParameters
    ig = sc.in_grid
    g = RadarGridParameters(ig.sensing_start - 60 / ig.prf, ig.wavelength, ig.prf / 4,
                            ig.starting_range - 10 * ig.range_pixel_spacing,
                            ig.range_pixel_spacing * 16, ig.lookside, 120, 36, ig.ref_epoch)
    z, h, err = gpu(sc, raw, out_grid=g, precision=precision)
    ref, _, rc = orc(sc, raw, out_grid=g, mode=0 if precision == "plain" else 1)
    assert err == (rc == 6)
    assert (ref == 0).any(), "some lines should have an empty aperture"
    np.testing.assert_array_equal(z == 0, ref == 0)
    tol = 1e-5 if precision == "plain" else 1e-3
    assert complex_rel_error(z, ref) <= tol


@pytest.mark.parametrize("precision", ["plain", "split"])
def test_resampled_output_grid(torch_gpu, precision):
    sc, raw = scene_and_raw("small")
    g = sc.out_grid.offset_and_resize(10.3, 20.7, 77, 53)
    g.prf = sc.out_grid.prf * 1.37
    g.range_pixel_spacing = sc.out_grid.range_pixel_spacing * 0.61
    z, _, _ = gpu(sc, raw, out_grid=g, precision=precision)
    ref, _, _ = orc(sc, raw, out_grid=g, mode=0 if precision == "plain" else 1)
    tol = 1e-5 if precision == "plain" else 1e-3
    assert complex_rel_error(z, ref) <= tol
    assert rms(phase_errors(z, ref)) <= tol


def test_empty_output_grid(torch_gpu):
    sc, raw = scene_and_raw("tiny")
    g = sc.out_grid.offset_and_resize(0, 0, 0, 5)
    z, h, err = gpu(sc, raw, out_grid=g)
    assert z.shape == (0, 5) and not err


def test_small_scene_golden_fixture(torch_gpu):
    fx = np.load(os.path.join(GOLDEN, "small_oracle.npz"))
    sc, raw = scene_and_raw("small")
    stats = {}
    z, h, err = gpu(sc, raw, stats=stats)
    assert not err and stats["updates"] == int(fx["updates"])
    idx = fx["index"]
    assert complex_rel_error(z.ravel()[idx], fx["out"]) <= 1e-5
    np.testing.assert_allclose(h.ravel()[idx], fx["height"], atol=1e-3)


@pytest.mark.parametrize("precision", ["plain", "split"])
def test_point_target_quality(torch_gpu, precision):
    """backproject.py:160-170 thresholds on the GPU output: position error below
    resolution/128, -3 dB widths within the range resolution and 1.103 ds."""
    sc, raw = scene_and_raw("pt")
    z, h, err = gpu(sc, raw, precision=precision)
    assert not err
    np.testing.assert_allclose(h, sc.dem.ref_height, atol=1e-6)
    g = sc.out_grid
    t_tgt, r_tgt = sc.target_radar[0]
    row = (t_tgt - g.sensing_start) * g.prf
    col = (r_tgt - g.starting_range) / g.range_pixel_spacing
    r = g.starting_range + g.range_pixel_spacing * np.arange(g.width)
    info = analyze_point_target(z * np.exp(-1j * 4 * np.pi / g.wavelength * r)[None, :], row, col)
    dr, dt = g.range_pixel_spacing, 1.0 / g.prf
    range_res = C / (2 * 0.8 * C / (2 * dr))
    _, vel = sc.orbit.interpolate(t_tgt)
    speed = np.linalg.norm(vel)
    assert abs(dr * info["range_offset"]) < range_res / 128
    assert abs(dt * info["azimuth_offset"] * speed) < sc.ds / 128
    assert dr * info["range_width"] <= range_res
    assert dt * info["azimuth_width"] * speed <= 6.62 / 6.0 * sc.ds


@pytest.mark.parametrize("precision", ["plain", "split"])
def test_device_tensor_io_matches_host_arrays(torch_gpu, precision):
    torch = torch_gpu
    from isce3_amd.focus import backproject
    sc, raw = scene_and_raw("tiny")
    host, hh, _ = gpu(sc, raw, precision=precision)
    d_raw = torch.from_numpy(raw.copy()).cuda()
    d_out = torch.full(sc.out_grid.shape, 7 + 7j, dtype=torch.complex64, device="cuda")
    d_h = torch.zeros(sc.out_grid.shape, dtype=torch.float32, device="cuda")
    err = backproject(d_out, sc.out_geometry(), d_raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                      sc.kernel, sc.dry_tropo_model, height=d_h, precision=precision)
    assert not err
    np.testing.assert_array_equal(d_out.cpu().numpy(), host)
    np.testing.assert_array_equal(d_h.cpu().numpy(), hh)


def test_context_reuse_across_sizes(torch_gpu):
    """Consecutive calls of different sizes on one context give the results of
    fresh calls (scratch buffers are grown and reused, never stale)."""
    sc1, raw1 = scene_and_raw("small")
    sc2, raw2 = scene_and_raw("tiny")
    a1, _, _ = gpu(sc1, raw1, precision="split")
    b1, _, _ = gpu(sc2, raw2, precision="split")
    a2, _, _ = gpu(sc1, raw1, precision="split")
    b2, _, _ = gpu(sc2, raw2, precision="split")
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(b1, b2)


@pytest.mark.parametrize("shape", [(1, 1), (3, 65), (17, 129), (9, 200)])
@pytest.mark.parametrize("precision", ["plain", "split"])
def test_ragged_output_grids(torch_gpu, shape, precision):
    """Output grids whose tile counts are not multiples of the workgroup's
    wavefront count leave wavefronts past the grid; they must do nothing."""
    sc, raw = scene_and_raw("tiny")
    g = sc.out_grid.offset_and_resize(3.0, 2.0, shape[0], shape[1])
    g.range_pixel_spacing = sc.out_grid.range_pixel_spacing * 0.37
    z, _, err = gpu(sc, raw, out_grid=g, precision=precision)
    ref, _, rc = orc(sc, raw, out_grid=g, mode=0 if precision == "plain" else 1)
    tol = 1e-5 if precision == "plain" else 1e-3
    assert complex_rel_error(z, ref) <= tol


@pytest.mark.parametrize("precision", ["plain", "split"])
def test_lines_narrower_than_the_kernel(torch_gpu, precision):
    """Range lines of 5 samples (fewer than the 8 taps): every tap goes through
    the zero-padded path, as in core/detail/Interp1d.h:62-83."""
    from isce3_amd.container import RadarGeometry
    sc, raw = scene_and_raw("tiny")
    ig = sc.in_grid.copy()
    ig.width = 5
    ig.starting_range = sc.out_grid.starting_range + 3 * sc.out_grid.range_pixel_spacing
    narrow = np.ascontiguousarray(raw[:, 100:105])
    geom = RadarGeometry(ig, sc.orbit, sc.doppler)
    g = sc.out_grid.offset_and_resize(0, 0, 16, 12)
    z, _, _ = gpu(sc, narrow, out_grid=g, precision=precision, in_geom=geom)
    ref, _, _ = orc(sc, narrow, out_grid=g, mode=0 if precision == "plain" else 1, in_geom=geom)
    assert np.abs(ref).max() > 0
    tol = 1e-5 if precision == "plain" else 1e-3
    assert complex_rel_error(z, ref) <= tol


def _dem_variant(sc, method, epsg):
    """The scene's DEM samples re-posted in projection `epsg` (30 m postings
    for projected codes) around the same centre, interpolated by `method`."""
    import math
    from isce3_amd.core import make_projection
    from isce3_amd.geometry import DEMInterpolator
    d = sc.dem
    lon_c = math.radians(d.x_start + 0.5 * d.width * d.delta_x)
    lat_c = math.radians(d.y_start + 0.5 * d.length * d.delta_y)
    xc, yc, _ = make_projection(epsg).forward([lon_c, lat_c, 0.0])
    dx, dy = (d.delta_x, d.delta_y) if epsg == 4326 else (30.0, -30.0)
    out = DEMInterpolator(d.ref_height, method=method)
    out.load_dem_array(d.data, xc - 0.5 * d.width * dx, yc - 0.5 * d.length * dy, dx, dy, epsg)
    return out


@pytest.mark.parametrize("method,epsg", [("biquintic", 4326), ("bicubic", 4326),
                                         ("nearest", 4326), ("biquintic", 32631),
                                         ("bilinear", 32631), ("bicubic", 6933),
                                         ("biquintic", 4978), ("bilinear", 3031),
                                         ("biquintic", 3413), ("sinc", 4326), ("sinc", 32631)])
@pytest.mark.parametrize("precision", ["plain", "split"])
def test_dem_methods_and_projections(torch_gpu, method, epsg, precision):
    """DEM rasters in UTM / EASE-Grid / geocentric coordinates and the
    nearest / bicubic / biquintic / sinc interpolators (geometry/DEMInterpolator.cpp,
    core/Projections.cpp, core/*Interpolator.cpp) against the oracle."""
    import dataclasses
    sc, raw = scene_and_raw("tiny")
    sc2 = dataclasses.replace(sc, dem=_dem_variant(sc, method, epsg))
    print(f"[dem] method={method} epsg={epsg} precision={precision}", flush=True)
    z, h, err = gpu(sc2, raw, precision=precision)
    ref, href, rc = orc(sc2, raw, mode=0 if precision == "plain" else 1)
    assert err == (rc == 6)
    if precision == "plain":
        assert_plain_parity(z, ref)
    else:
        assert complex_rel_error(z, ref) <= 1e-3
        assert rms(phase_errors(z, ref)) <= 1e-3
    if method == "sinc":
        # the sinc interpolant takes the nearest of 8192 fractional kernel rows
        # (Sinc2dInterpolator.cpp:77-79), so the DEM is piecewise constant on
        # steps of 1/8192 posting: FP64 coordinates that differ in the last
        # bits between GPU and CPU put a few Brent roots on the neighbouring
        # step, millimetres apart on this fractal terrain
        dh = np.abs(h - href)
        assert np.mean(dh <= 1e-3) >= 0.99 and dh.max() <= 2e-2, (np.mean(dh <= 1e-3), dh.max())
    else:
        np.testing.assert_allclose(h, href, atol=1e-3)
    # the DEM is in play: heights differ from the flat reference height
    assert np.nanstd(h) > 1.0
