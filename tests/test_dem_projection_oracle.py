"""DEM generality on the host: map projections and 2-D interpolators.

The oracle's restatement (oracle/bp_oracle.cpp) is pinned to the reference's
own known answers -- 308 projection vectors of tests/cxx/isce3/core/
projections/{utm,polar,cea,geocent}.cpp (collected by
tests/golden/make_projection_kats.py) at the reference's tolerances, and the
interpolator test data tests/data/interpolator/data.txt with the thresholds of
tests/cxx/isce3/core/interpolator/interpolator.cpp:100-190 -- plus exactness
properties (node values, planes).  The product's host code (the same FP64
routines the geometry kernel runs) is then checked against the oracle."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "projection_kats.json")))
SINC, BILINEAR, BICUBIC, NEAREST, BIQUINTIC = 0, 1, 2, 3, 4


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def orc_forward(epsg, llh):
    out = (C.c_double * 3)()
    rc = oracle.lib().orc_proj_forward(epsg, _d3(llh), out)
    return rc, list(out)


def orc_inverse(epsg, xyz):
    out = (C.c_double * 3)()
    rc = oracle.lib().orc_proj_inverse(epsg, _d3(xyz), out)
    return rc, list(out)


def test_kat_fixture_covers_every_projection():
    codes = {k["epsg"] for k in KATS}
    assert {4978, 3031, 3413, 6933} <= codes
    assert sum(1 for c in codes if 32600 < c <= 32660) == 60
    assert sum(1 for c in codes if 32700 < c <= 32760) == 60
    assert len(KATS) == 308


def test_oracle_projections_match_reference_known_answers():
    """projtest.h:15-26: forward within 1e-6 m, inverse within 1e-9 rad and
    1e-6 m."""
    for k in KATS:
        rc, xyz = orc_forward(k["epsg"], k["llh"])
        assert rc == 0
        assert np.allclose(xyz, k["xyz"], rtol=0, atol=1e-6), (k, xyz)
        rc, llh = orc_inverse(k["epsg"], k["xyz"])
        assert rc == 0
        assert abs(llh[0] - k["llh"][0]) < 1e-9 and abs(llh[1] - k["llh"][1]) < 1e-9, (k, llh)
        assert abs(llh[2] - k["llh"][2]) < 1e-6
    assert orc_forward(4327, [0, 0, 0])[0] == -1
    assert orc_forward(32661, [0, 0, 0])[0] == -1


def test_product_projections_match_known_answers_and_oracle():
    from isce3_amd import core, errors
    for k in KATS:
        p = core.make_projection(k["epsg"])
        xyz = p.forward(k["llh"])
        assert np.allclose(xyz, k["xyz"], rtol=0, atol=1e-6)
        llh = p.inverse(k["xyz"])
        assert np.allclose(llh[:2], k["llh"][:2], rtol=0, atol=1e-9)
        _, oxyz = orc_forward(k["epsg"], k["llh"])
        assert np.allclose(xyz, oxyz, rtol=0, atol=1e-8)
    assert repr(core.UTM(32611)) == "UTM(epsg=32611)"
    assert repr(core.make_projection(3031)) == "PolarStereo(epsg=3031)"
    assert core.CEA().code == 6933 and core.LonLat().code == 4326
    with pytest.raises(RuntimeError):
        core.make_projection(1234)
    with pytest.raises(ValueError):
        core.UTM(32661)
    # beyond the UTM easting limit the forward transform reports failure
    with pytest.raises(errors.RuntimeError_):
        core.UTM(32631).forward([math.radians(92.9), 0.0, 0.0])


# ------------------------------------------------------------ interpolators
def _arange(lo, hi, inc):
    out, c = [], lo
    while c < hi:
        out.append(c)
        c += inc
    return np.array(out)


def _reference_grid():
    # interpolator.cpp:60-90: z = sin(x^2 + y^2) on arange(-5.01, 5.01, 0.25)
    xi = _arange(-5.01, 5.01, 0.25)
    yi = _arange(-5.01, 5.01, 0.25)
    M = np.sin(yi[:, None] ** 2 + xi[None, :] ** 2)
    data = np.loadtxt(os.path.join(HERE, "golden", "interpolator_data.txt"))
    return M, data, -5.01, 0.25


def _interp(method, x, y, z):
    z = np.ascontiguousarray(z)
    if z.dtype == np.float64:
        return oracle.lib().orc_interp2d_f64(method, x, y, z.ctypes.data, z.shape[1], z.shape[0])
    return oracle.lib().orc_interp2d_f32(method, x, y, z.ctypes.data, z.shape[1], z.shape[0])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_interpolators_meet_reference_tests(dtype):
    M, data, start, delta = _reference_grid()
    Mz = M.astype(dtype)
    n = len(data)
    err = {BILINEAR: 0.0, BICUBIC: 0.0, BIQUINTIC: 0.0, SINC: 0.0}
    n_sinc = 0
    tol = 1e-8 if dtype == np.float64 else 2e-6
    for row in data:
        x, y = (row[0] - start) / delta, (row[1] - start) / delta
        # Nearest (interpolator.cpp:100-118)
        assert _interp(NEAREST, x, y, Mz) == pytest.approx(
            float(Mz[int(round(y)), int(round(x))]), abs=1e-8)
        # Bilinear against the data's own bilinear column (:121-141)
        zb = _interp(BILINEAR, x, y, Mz)
        assert abs(zb - row[2]) < tol
        err[BILINEAR] += (zb - row[5]) ** 2
        # Bicubic within the interior (:144-168)
        if x >= 1 and y >= 1 and x < M.shape[1] - 2 and y < M.shape[0] - 2:
            err[BICUBIC] += (_interp(BICUBIC, x, y, Mz) - row[5]) ** 2
        err[BIQUINTIC] += (_interp(BIQUINTIC, x, y, Mz) - row[5]) ** 2
        # Sinc (8 taps, 8192 fractional rows) where its window fits (:210-236)
        if x >= 3 and y >= 3 and x <= M.shape[1] - 5 and y <= M.shape[0] - 5:
            err[SINC] += (_interp(SINC, x, y, Mz) - row[5]) ** 2
            n_sinc += 1
    assert err[BILINEAR] / n < 0.07
    assert err[BICUBIC] / n < 0.058
    assert err[BIQUINTIC] / n < 0.058
    assert n_sinc > n // 2 and err[SINC] / n_sinc < 0.0003


def test_oracle_interpolators_exact_at_nodes_and_on_planes():
    rng = np.random.default_rng(0)
    z = rng.standard_normal((20, 24)).astype(np.float32)
    for method in (BILINEAR, BICUBIC, NEAREST, BIQUINTIC):
        for r, c in ((2, 2), (7, 11), (17, 21)):
            assert _interp(method, float(c), float(r), z) == z[r, c]
    # sinc: the kernel row of fraction 0 is one unit tap (sin(pi n) / (pi n)
    # of the other taps is ~1e-16), so float nodes come back exactly; outside
    # the 8 x 8 window the reference returns zero; a constant field stays
    # constant (every row of the kernel sums to one)
    for r, c in ((3, 3), (7, 11), (15, 19)):
        assert _interp(SINC, float(c), float(r), z) == z[r, c]
    assert _interp(SINC, 2.5, 8.0, z) == 0.0 and _interp(SINC, 10.0, 16.5, z) == 0.0
    assert _interp(SINC, 10.0, 15.5, z) != 0.0   # row 15 is the last whose window fits
    const = np.full((20, 24), 7.25, np.float64)
    for x, y in ((3.3, 4.7), (10.01, 12.99), (19.0, 15.0)):
        assert _interp(SINC, x, y, const) == pytest.approx(7.25, abs=1e-12)
    yy, xx = np.mgrid[0:20, 0:24]
    plane = (3.0 + 0.25 * xx - 0.5 * yy).astype(np.float64)
    for method in (BILINEAR, BICUBIC, BIQUINTIC):
        for x, y in ((3.3, 4.7), (10.01, 12.99), (18.5, 13.25)):   # no edge clamping
            assert _interp(method, x, y, plane) == pytest.approx(3.0 + 0.25 * x - 0.5 * y,
                                                                 abs=1e-12)


@pytest.mark.parametrize("method", ["bilinear", "bicubic", "nearest", "biquintic", "sinc"])
@pytest.mark.parametrize("epsg", [4326, 32631, 3413, 6933])
def test_product_dem_lookup_matches_oracle(method, epsg):
    """DEMInterpolator.interpolate_lonlat / interpolate_xy of the product's
    host code against the oracle, for each method and projection."""
    from isce3_amd import core
    from isce3_amd.geometry import DEMInterpolator
    rng = np.random.default_rng(epsg + len(method))
    z = (rng.standard_normal((64, 80)) * 50 + 300).astype(np.float32)
    # a raster around (lon, lat) = (2 deg, 70 deg) in the projection's coordinates
    lon0, lat0 = math.radians(2.0), math.radians(70.0)
    p = core.make_projection(epsg)
    x0, y0, _ = p.forward([lon0, lat0, 0.0])
    dx, dy = (0.001, -0.001) if epsg == 4326 else (90.0, -90.0)
    dem = DEMInterpolator(10.0, method=method)
    dem.load_dem_array(z, x0 - 40 * dx, y0 - 32 * dy, dx, dy, epsg)
    od = oracle.dem_struct(dem)
    for k in range(200):
        lon = lon0 + rng.uniform(-1, 1) * (0.03 if epsg == 4326 else 0.08)
        lat = lat0 + rng.uniform(-1, 1) * 0.02
        a = dem.interpolate_lonlat(lon, lat)
        b = oracle.lib().orc_dem_interpolate_lonlat(C.byref(od), lon, lat)
        # host code on both sides rounds identically (no fused multiply-add)
        assert a == b, (method, epsg, lon, lat)
        xs, ys, _ = p.forward([lon, lat, 0.0])
        assert dem.interpolate_xy(xs, ys) == oracle.lib().orc_dem_interpolate_xy(
            C.byref(od), xs, ys)
