"""Pins of the RSLC block-finishing oracle (oracle/rslc.py): the reference's
own statistics tests (tests/python/extensions/pybind/math/stats.py:28-176),
IEEE half-precision facts for the complex32 encoding, bit patterns of the
mantissa truncation, and the complex product's rounding."""
import math

import numpy as np
import numpy.testing as npt
import pytest

import rslc as ro


def _analytic_arrays(length=10, width=10):
    # stats.py:46-56
    i = np.arange(length, dtype=np.float64)[:, None]
    j = np.arange(width, dtype=np.float64)[None, :]
    re = np.sin(2. * np.pi * i / length) + np.cos(4. * np.pi * j / width)
    im = 2. * (np.sin(6. * np.pi * i / length) - np.cos(8. * np.pi * j / width))
    return re, im


def test_welford_matches_reference_stats_test_run():
    """stats.py:28-121: min/mean/max/sample stddev of the float32 rasters
    against numpy at atol 1e-7."""
    re, im = _analytic_arrays()
    z = (re.astype(np.float32) + 1j * im.astype(np.float32)).astype(np.complex64)
    sre, sim = ro.stats_welford(z)
    re32, im32 = re.astype(np.float32), im.astype(np.float32)
    for s, a in ((sre, re32), (sim, im32)):
        npt.assert_allclose(s.min, a.min(), atol=1e-7)
        npt.assert_allclose(s.max, a.max(), atol=1e-7)
        npt.assert_allclose(float(s.mean), a.astype(np.float64).mean(), atol=1e-7)
        npt.assert_allclose(s.sample_stddev, a.astype(np.float64).std(ddof=1), atol=1e-7)
        assert s.n_valid == 100


def test_welford_matches_reference_stats_test_array():
    """stats.py:124-160: seeded normal arrays, rtol 1e-6; merging a copy
    doubles n_valid and keeps min/max/mean."""
    np.random.seed(0)
    x = np.random.normal(size=(5, 3)).astype("f4")
    y = np.random.normal(size=(5, 3)).astype("f4")
    s, _ = ro.stats_welford(x + 0j)
    npt.assert_allclose(float(s.mean), x.mean(), rtol=1e-6)
    npt.assert_allclose(s.min, x.min(), rtol=1e-6)
    npt.assert_allclose(s.max, x.max(), rtol=1e-6)
    npt.assert_allclose(s.sample_stddev, x.std(ddof=1), rtol=1e-6)
    s3, _ = ro.stats_welford(x + 0j)
    s3.update(s)
    assert s3.n_valid == 2 * s.n_valid
    npt.assert_allclose(float(s3.mean), float(s.mean), rtol=1e-6)
    _, si = ro.stats_welford((x + 1j * y).astype(np.complex64))
    npt.assert_allclose(si.sample_stddev, y.std(ddof=1), rtol=1e-6)


def test_ones_give_zero_stddev():
    """stats.py:163-176: ten ones -> n_valid 10, mean/min/max 1, stddev 0."""
    s, si = ro.stats_welford(np.ones(10, np.float32) * (1 + 1j))
    assert (s.n_valid, float(s.mean), s.min, s.max, s.sample_stddev) == (10, 1.0, 1.0, 1.0, 0.0)
    assert (si.n_valid, float(si.mean)) == (10, 1.0)


def test_exact_and_welford_agree_and_chan_merge_is_exact():
    rng = np.random.default_rng(4)
    z = (rng.standard_normal(3001) * 3 + 1 + 1j * (rng.standard_normal(3001) - 2)).astype(np.complex64)
    z[[5, 77, 901]] = np.nan + 0j
    z[[6, 1000]] = 1 + 1j * np.nan
    sre, sim = ro.stats_welford(z)
    ex = ro.stats_exact(z)
    assert sre.n_valid == sim.n_valid == ex["n_valid"] == 3001 - 5
    for s, (lo, hi, m, m2) in ((sre, ex["real"]), (sim, ex["imag"])):
        assert (s.min, s.max) == (lo, hi)
        assert abs(s.real_valued_mean - m) < 1e-12 * (1 + abs(m))
        assert abs(s.square_diff_sum - m2) < 1e-9 * m2
        assert abs(float(s.mean) - m) < 1e-5
    # Chan's merge of two halves equals the whole
    a, _ = ro.stats_welford(z[:1400])
    b, _ = ro.stats_welford(z[1400:])
    a.update(b)
    assert a.n_valid == sre.n_valid
    assert abs(a.real_valued_mean - sre.real_valued_mean) < 1e-12
    assert abs(a.square_diff_sum - sre.square_diff_sum) < 1e-9 * sre.square_diff_sum


def test_complex32_encoding_ieee_half_facts():
    vals = np.array([1.0, 65504.0, 65520.0, 2.0 ** -24, 1.0 / 3.0, 1 + 2.0 ** -11,
                     1 + 3 * 2.0 ** -11, -2.0, 2.0 ** -25, 2.0 ** -25 * 1.5, 1e6, -0.0],
                    np.float32)
    want = [0x3C00, 0x7BFF, 0x7C00, 0x0001, 0x3555, 0x3C00, 0x3C02, 0xC000, 0x0000, 0x0001,
            0x7C00, 0x8000]
    z = np.zeros(vals.shape, np.complex64)
    z.real = vals
    enc = ro.to_complex32(z)
    got = enc["r"].view(np.uint16)
    assert [int(g) for g in got] == want
    assert np.all(enc["i"].view(np.uint16)[:-1] == 0)


def test_truncate_mantissa_bit_patterns():
    z = np.array([np.pi + 1j * -np.e], np.complex64)
    t = ro.truncate_mantissa(z, 10)
    u = t.view(np.uint32)
    assert int(u[0]) == 0x40490FDB & ~0x1FFF
    assert int(u[1]) == int(np.array([-np.e], np.float32).view(np.uint32)[0]) & ~0x1FFF
    assert np.array_equal(ro.truncate_mantissa(z, 23).view(np.uint32), z.view(np.uint32))
    assert int(ro.truncate_mantissa(z, 1).view(np.uint32)[0]) == 0x40400000
    for bad in (0, 24, -3):
        with pytest.raises(ValueError):
            ro.truncate_mantissa(z, bad)


def test_rounded_product_is_within_an_ulp_of_the_exact_product():
    rng = np.random.default_rng(9)
    z = (rng.standard_normal(5000) + 1j * rng.standard_normal(5000)).astype(np.complex64)
    c = np.exp(1j * rng.uniform(-np.pi, np.pi, 5000)).astype(np.complex64)
    p = ro.cmul_rounded(z, c)
    exact = z.astype(np.complex128) * c.astype(np.complex128)
    scale = np.abs(z.astype(np.complex128)) * np.abs(c.astype(np.complex128))
    err = np.abs(p.astype(np.complex128) - exact) / scale
    assert err.max() < 4 * 2.0 ** -24
    # numpy's own complex64 product (FMA or not, per build) agrees to the same bound
    assert (np.abs((z * c).astype(np.complex128) - p) / scale).max() < 4 * 2.0 ** -24
    # term-by-term rounding: real part is round(round(ac) - round(bd))
    ac = (z.real * c.real).astype(np.float32)
    bd = (z.imag * c.imag).astype(np.float32)
    assert np.array_equal(p.real, (ac - bd).astype(np.float32))


def test_finish_block_order_of_operations():
    rng = np.random.default_rng(2)
    z = (rng.standard_normal((3, 5)) + 1j * rng.standard_normal((3, 5))).astype(np.complex64)
    rc = np.exp(1j * rng.uniform(0, 6, 5)).astype(np.complex64)
    prior = (rng.standard_normal((3, 5)) + 0j).astype(np.complex64)
    zo, enc, st = ro.finish_block(z, rc, prior, "complex64_zero_mantissa", 5)
    scaled = ro.cmul_rounded(z, rc[None, :])
    acc = (scaled.real + prior.real) + 1j * (scaled.imag + prior.imag)
    ex = ro.stats_exact(acc.astype(np.complex64))
    assert st == ex       # statistics are of the values before truncation
    assert np.array_equal(zo.view(np.uint32),
                          ro.truncate_mantissa(acc.astype(np.complex64), 5).view(np.uint32))
    assert enc is zo
    assert math.isnan(ro.stats_exact(np.zeros(0, np.complex64))["real"][0])
