"""RSLC writer boundary checks that need no GPU: BackgroundWriter's
data-type and dataset-type rules (focus.py:845-872) and the finisher's
argument rules, raised before any device work."""
import numpy as np
import pytest

from isce3_amd import errors, rslc


class FakeDataset:
    def __init__(self, shape, dtype):
        self.data = np.zeros(shape, dtype)

    @property
    def dtype(self):
        return self.data.dtype

    def __getitem__(self, key):
        return self.data[key]

    def write_direct(self, a, dest_sel):
        self.data[dest_sel] = a


def test_writer_type_rules():
    rc = np.ones(8, np.complex64)
    with pytest.raises(ValueError):
        rslc.BackgroundWriter(rc, FakeDataset((4, 8), np.complex64), "complex128")
    with pytest.raises(TypeError):
        rslc.BackgroundWriter(rc, FakeDataset((4, 8), np.complex64), "complex32")
    with pytest.raises(TypeError):
        rslc.BackgroundWriter(rc, FakeDataset((4, 8), rslc.complex32), "complex64")
    with pytest.raises(ValueError):
        rslc.BackgroundWriter(rc, FakeDataset((4, 8), np.complex64), "complex64_zero_mantissa",
                              mantissa_nbits=0)
    w = rslc.BackgroundWriter(rc, FakeDataset((4, 8), rslc.complex32), "complex32")
    w.notify_finished()
    with pytest.raises(RuntimeError):
        w.queue_write(np.zeros((4, 8), np.complex64), np.s_[0:4, 0:8])


def test_truncate_mantissa_argument_rules():
    with pytest.raises(ValueError):
        rslc.truncate_mantissa([1.0, 2.0])
    with pytest.raises(ValueError):
        rslc.truncate_mantissa(np.zeros(3, np.complex64), 0)
    z = np.ones(3, np.complex64)
    rslc.truncate_mantissa(z, 23)            # no-op, no device work
    assert np.all(z == 1)


def test_stats_records_merge_and_print():
    a = rslc.StatsFloat32()
    assert str(a) == "StatsFloat32(n_valid=0, mean=0.0, min=nan, max=nan, sample_stddev=0.0)"
    b = rslc.StatsFloat32()
    b.min, b.max, b.mean, b.n_valid, b._mean64, b._m2 = 1.0, 1.0, 1.0, 10, 1.0, 0.0
    assert str(b) == "StatsFloat32(n_valid=10, mean=1.0, min=1.0, max=1.0, sample_stddev=0.0)"
    a.update(b)
    a.update(b)
    assert (a.n_valid, a.mean, a.min, a.max, a.sample_stddev) == (20, 1.0, 1.0, 1.0, 0.0)
    c = rslc.StatsRealImagFloat32()
    c.real, c.imag, c.n_valid = b, b, 10
    assert str(c) == (f"StatsRealImagFloat32(real={b}, imag={b})")
