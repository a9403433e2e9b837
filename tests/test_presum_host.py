"""Presumming without a GPU: the oracle (oracle/presum.py) and the library's
host BLU weights against the reference's own tests
(tests/cxx/isce3/focus/presum.cpp, tests/python/extensions/pybind/focus/
presum.py) and the interpolation property of BLU weights."""
import numpy as np
import numpy.testing as npt
import pytest

import presum as po
from isce3_amd import core, errors
from isce3_amd.focus import (compute_ids_from_mask, fill_weights, get_presum_weights,
                             get_unique_ids)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_domain_and_values(impl):
    """presum.cpp:6-58: Bartlett(3) on t = 0..9."""
    t = np.arange(10.0)
    if impl == "oracle":
        f = lambda tout: po.presum_weights(po.bartlett(3.0), t, tout)  # noqa: E731
    else:
        f = lambda tout: get_presum_weights(core.BartlettKernel(3.0), t, tout)  # noqa: E731
    off, w = f(5.0)
    assert (len(w), off) == (3, 4)
    assert w[0] == pytest.approx(0.0, abs=4e-16) and w[1] == pytest.approx(1.0, abs=4e-16)
    assert w[2] == pytest.approx(0.0, abs=4e-16)
    off, w = f(5.5)
    assert (len(w), off) == (4, 4)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_no_data_in_a_gap(impl):
    """presum.cpp:61-76: AzimuthKernel(1), samples at 0 and 10, output at 5."""
    t = np.array([0.0, 10.0])
    off, w = (po.presum_weights(po.azimuth(1.0), t, 5.0) if impl == "oracle"
              else get_presum_weights(core.AzimuthKernel(1.0), t, 5.0))
    assert len(w) == 0 and 0 <= off < len(t)


def test_weights_window_and_agreement_with_oracle():
    """pybind presum.py:7-17 (window within L), and the pivoted LDL^T of the
    library against numpy's LU on jittered sample times."""
    n = 31
    rng = np.random.default_rng(12345)
    t = np.linspace(0, 9, n) + rng.normal(0.1, size=n)
    t.sort()
    L = 1.0
    for tout in np.linspace(0.5, 8.5, 17):
        off, w = get_presum_weights(core.AzimuthKernel(L), t, tout)
        i = slice(off, off + len(w))
        assert all(abs(t[i] - tout) <= L)
        ooff, ow = po.presum_weights(po.azimuth(L), t, tout)
        assert off == ooff and len(w) == len(ow)
        npt.assert_allclose(w, ow, rtol=1e-9, atol=1e-9)


def test_blu_weights_interpolate_at_sample_times():
    """At an input time the right-hand side is a column of the system, so
    the weights are the unit vector of that sample."""
    rng = np.random.default_rng(7)
    t = np.sort(np.arange(40) * 0.66e-3 + rng.uniform(-0.1e-3, 0.1e-3, 40))
    for k in (5, 17, 30):
        for impl in ("oracle", "product"):
            off, w = (po.presum_weights(po.azimuth(1.6e-3), t, t[k]) if impl == "oracle"
                      else get_presum_weights(core.AzimuthKernel(1.6e-3), t, t[k]))
            e = np.zeros(len(w))
            e[k - off] = 1.0
            npt.assert_allclose(w, e, atol=1e-9)


def test_kernels_closed_forms():
    az = core.AzimuthKernel(2.0)          # width 4, x = |t|
    assert az.width == 4.0
    for x, v in ((0.0, 1.0), (0.5, 0.25 * 0.75 * 0.5 - 0.375 + 1.0), (1.0, 0.25),
                 (1.5, 0.25 * 0.125), (2.0, 0.0), (2.5, 0.0)):
        assert az(x) == pytest.approx(v) and az(-x) == pytest.approx(v)
    b = core.BartlettKernel(3.0)
    assert (b(0.0), b(0.75), b(1.5), b(2.0)) == (1.0, 0.5, 0.0, 0.0)
    f, _ = po.azimuth(2.0)
    xs = np.linspace(-2.2, 2.2, 45)
    npt.assert_allclose([az(x) for x in xs], f(xs), atol=1e-15)


def test_fill_weights_ids_and_unique():
    """pybind presum.py:20-85."""
    lut = {123: np.array([1, 2, 3.]), 456: np.array([4, 5, 6.]), 789: np.array([7, 8, 9.])}
    ids = np.random.default_rng(0).choice(list(lut.keys()), 1000)
    w = fill_weights(ids, lut)
    assert w.shape == (3, 1000)
    for i in (0, 10, 999):
        assert np.all(w[:, i] == lut[ids[i]])
    with pytest.raises(errors.LengthError):
        fill_weights(ids, {})
    mask = np.zeros((499, 4), dtype=bool)
    for j in range(4):
        mask[j:, j] = True
    expected = np.zeros(499, np.int64)
    expected[0], expected[1], expected[2], expected[3:] = 1, 3, 7, 15
    npt.assert_array_equal(compute_ids_from_mask(mask), expected)
    npt.assert_array_equal(po.compute_ids_from_mask(mask), expected)
    big = np.zeros((3, 63), bool)
    big[:, 62] = True
    assert compute_ids_from_mask(big)[0] == 1 << 62
    with pytest.raises(errors.LengthError):
        compute_ids_from_mask(np.zeros((3, 64), bool))
    ids = np.array([1, 1, 1, 3, 3, 3, 2, 2, 0, 0, 0, 0, 0, 0])
    npt.assert_array_equal(get_unique_ids(ids), np.unique(ids))


def test_oracle_apply_weights_formula():
    t = np.array([-2.2, -1.1, -0.1, 0.9, 1.9]) / 1910.
    fd = np.linspace(800, 900, 7)
    rng = np.random.default_rng(1)
    w = rng.normal(size=(5, 7))
    x = rng.normal(size=(5, 7)) + 1j * rng.normal(size=(5, 7))
    out = po.apply_presum_weights(t, fd, w, x)
    j = 3
    want = sum(w[i, j] * np.exp(-2j * np.pi * t[i] * fd[j]) * x[i, j] for i in range(5))
    assert abs(out[j] - want) < 1e-12


def test_host_allocation_failure_is_an_error_not_an_abort():
    """A window of 2^23 samples asks the weight solve for a 2^23 x 2^23
    system (512 TiB, beyond any process's address space): the C-ABI call
    returns BP_ERR_OUT_OF_MEMORY (MemoryError here) instead of letting the
    C++ exception end the process, and the next call works."""
    t = np.arange(1 << 23) * 1e-3
    with pytest.raises(MemoryError, match="host memory allocation failed"):
        get_presum_weights(core.BartlettKernel(1e5), t, 10.0)
    off, w = get_presum_weights(core.BartlettKernel(2.5e-3), t[:40], 10e-3)
    assert off == 9 and len(w) == 3
    npt.assert_allclose(w, [0.0, 1.0, 0.0], atol=1e-12)
