"""RSLC block finishing on the GPU (include/isce3_amd_rslc.h through
isce3_amd.rslc) against the oracle (oracle/rslc.py, a restatement of
nisar/workflows/focus.py:883-914 and isce3/math/Stats.cpp).

Bar: scaled/accumulated samples, complex32 encodings and truncated mantissas
bit-identical to the oracle; n_valid, min and max exact; FP64 mean within
1e-12 and sum of squared deviations within 1e-10 relative of the oracle's
two-pass FP64 values (different summation order); the reference's own
statistics tests (stats.py) at their tolerances."""
import math

import numpy as np
import numpy.testing as npt
import pytest

import rslc as ro

pytestmark = pytest.mark.gpu


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


def random_block(rng, rows, cols, scale=1.0):
    return ((rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))) *
            scale).astype(np.complex64)


def check_stats(s, ex):
    assert s.n_valid == ex["n_valid"]
    for part, (lo, hi, m, m2) in ((s.real, ex["real"]), (s.imag, ex["imag"])):
        assert part.n_valid == ex["n_valid"]
        if ex["n_valid"] == 0:
            assert math.isnan(part.min) and math.isnan(part.max)
            continue
        assert (part.min, part.max) == (lo, hi)
        scale = max(abs(lo), abs(hi), 1e-30)
        assert abs(part._mean64 - m) <= 1e-12 * scale
        assert abs(part._m2 - m2) <= 1e-10 * max(m2, 1e-30)
        assert part.mean == float(np.float32(part._mean64))
        if ex["n_valid"] > 1:
            npt.assert_allclose(part.sample_stddev, math.sqrt(m2 / (ex["n_valid"] - 1)),
                                rtol=1e-6)


SHAPES = [(1, 1), (1, 7), (3, 5), (7, 1), (17, 255), (64, 256), (33, 1000), (257, 4097)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("data_type,nbits", [("complex64", 10), ("complex32", 10),
                                             ("complex64_zero_mantissa", 10),
                                             ("complex64_zero_mantissa", 1)])
def test_finish_block_matches_oracle(torch_gpu, shape, data_type, nbits):
    from isce3_amd.rslc import BlockFinisher
    rows, cols = shape
    rng = np.random.default_rng(rows * 7919 + cols + nbits)
    z = random_block(rng, rows, cols, 3.0)
    rc = np.exp(1j * rng.uniform(-np.pi, np.pi, cols)).astype(np.complex64)
    prior = random_block(rng, rows, cols) if rows * cols % 2 else None
    zo, enc_o, st_o = ro.finish_block(z, rc, prior, data_type, nbits)
    f = BlockFinisher(max(rows * cols, 4096))
    zg = z.copy()
    enc, st = f.finish(zg, range_cor=rc, prior=prior, data_type=data_type, mantissa_nbits=nbits)
    print(f"[rslc] shape={shape} type={data_type} nbits={nbits} n={st.n_valid}", flush=True)
    assert np.array_equal(zg.view(np.uint32), zo.view(np.uint32))
    if data_type == "complex32":
        assert enc.dtype == ro.complex32
        assert np.array_equal(enc.view(np.uint16), enc_o.view(np.uint16))
    else:
        assert enc is zg
    check_stats(st, st_o)
    f.close()


def test_nan_samples_are_skipped_pairwise(torch_gpu):
    from isce3_amd.rslc import BlockFinisher
    rng = np.random.default_rng(3)
    z = random_block(rng, 40, 300)
    z[3, 7] = complex(np.nan, 1.0)
    z[5, 299] = complex(2.0, np.nan)
    z[39, 0] = complex(np.nan, np.nan)
    z[0, :17] = complex(np.nan, 0.0)
    _, _, ex = ro.finish_block(z)
    _, st = BlockFinisher(z.size).finish(z.copy())
    assert st.n_valid == z.size - 20
    check_stats(st, ex)
    allnan = np.full((2, 3), np.nan, np.complex64)
    _, st = BlockFinisher(6).finish(allnan)
    assert st.n_valid == 0 and math.isnan(st.real.min) and st.real.sample_stddev == 0.0


def test_empty_block_and_limits(torch_gpu):
    from isce3_amd import errors
    from isce3_amd.rslc import BlockFinisher
    f = BlockFinisher(100)
    _, st = f.finish(np.zeros((0, 5), np.complex64))
    assert st.n_valid == 0 and math.isnan(st.imag.max)
    with pytest.raises(errors.LengthError):
        f.finish(np.zeros((11, 10), np.complex64))
    with pytest.raises(errors.InvalidArgument):
        f.finish(np.zeros((2, 2), np.complex64), data_type="complex64_zero_mantissa",
                 mantissa_nbits=24)
    with pytest.raises(errors.LengthError):
        f.finish(np.zeros((2, 2), np.complex64), range_cor=np.ones(3, np.complex64))


def test_reference_stats_tests_on_gpu(torch_gpu):
    """tests/python/extensions/pybind/math/stats.py:28-176 with the GPU
    computing every array's statistics."""
    from isce3_amd.rslc import StatsFloat32, StatsRealImagFloat32
    length = width = 10
    i = np.arange(length, dtype=np.float64)[:, None]
    j = np.arange(width, dtype=np.float64)[None, :]
    re = np.sin(2. * np.pi * i / length) + np.cos(4. * np.pi * j / width)
    im = 2. * (np.sin(6. * np.pi * i / length) - np.cos(8. * np.pi * j / width))
    sr = StatsFloat32(re.astype(np.float32))
    sri = StatsRealImagFloat32((re + 1j * im).astype(np.complex64))
    for s, a in ((sr, re), (sri.real, re), (sri.imag, im)):
        npt.assert_allclose(s.min, np.min(a), atol=1e-7)
        npt.assert_allclose(s.mean, np.mean(a), atol=1e-7)
        npt.assert_allclose(s.max, np.max(a), atol=1e-7)
        npt.assert_allclose(s.sample_stddev, np.std(a, ddof=1), atol=1e-7)
    np.random.seed(0)
    x = np.random.normal(size=(5, 3)).astype("f4")
    y = np.random.normal(size=(5, 3)).astype("f4")
    s = StatsFloat32(x)
    npt.assert_allclose(s.mean, x.mean(), rtol=1e-6)
    npt.assert_allclose(s.min, x.min(), rtol=1e-6)
    npt.assert_allclose(s.max, x.max(), rtol=1e-6)
    npt.assert_allclose(s.sample_stddev, x.std(ddof=1), rtol=1e-6)
    s2 = StatsFloat32()
    s2.update(x)
    assert (s2.min, s2.max, s2.n_valid) == (s.min, s.max, s.n_valid)
    npt.assert_allclose(s2.mean, s.mean, rtol=1e-6)
    s3 = StatsFloat32(x)
    s3.update(s)
    assert s3.n_valid == 2 * s.n_valid
    npt.assert_allclose(s3.mean, s.mean, rtol=1e-6)
    z = (x + 1j * y).astype("c8")
    sri = StatsRealImagFloat32(z)
    sri2 = StatsRealImagFloat32()
    sri2.update(z)
    sri2.update(sri)
    assert sri2.n_valid == 30
    ones = StatsFloat32()
    ones.update(np.ones(10).astype("f4"))
    expected = "StatsFloat32(n_valid=10, mean=1.0, min=1.0, max=1.0, sample_stddev=0.0)"
    assert str(ones) == expected
    cpx = StatsRealImagFloat32()
    cpx.update(np.ones(10).astype("f4") * (1 + 1j))
    assert str(cpx) == f"StatsRealImagFloat32(real={expected}, imag={expected})"


def test_complex32_and_truncation_special_values(torch_gpu):
    from isce3_amd.rslc import to_complex32, truncate_mantissa
    vals = np.array([1.0, 65504.0, 65520.0, 2.0 ** -24, 1.0 / 3.0, 1 + 2.0 ** -11,
                     1 + 3 * 2.0 ** -11, -2.0, 2.0 ** -25, 2.0 ** -25 * 1.5, 1e6, -0.0,
                     -1e-30, np.inf, -np.inf, 6.1e-5, 5.96e-8, 3.0e-6], np.float32)
    z = np.zeros(vals.shape, np.complex64)
    z.real = vals
    z.imag = vals[::-1]
    enc = to_complex32(z)
    want = ro.to_complex32(z)
    assert np.array_equal(enc.view(np.uint16), want.view(np.uint16))
    zn = np.array([complex(np.nan, 1.0)], np.complex64)
    e = to_complex32(zn)
    assert np.isnan(e["r"][0].astype(np.float32)) and e["i"][0] == 1.0
    zt = z.copy()
    truncate_mantissa(zt, 7)
    assert np.array_equal(zt.view(np.uint32), ro.truncate_mantissa(z, 7).view(np.uint32))


def test_device_tensor_in_place(torch_gpu):
    torch = torch_gpu
    from isce3_amd.rslc import BlockFinisher
    rng = np.random.default_rng(8)
    z = random_block(rng, 300, 700)
    rc = np.exp(1j * rng.uniform(0, 6, 700)).astype(np.complex64)
    zo, enc_o, ex = ro.finish_block(z, rc, None, "complex32")
    dz = torch.from_numpy(z).cuda()
    drc = torch.from_numpy(rc).cuda()
    enc, st = BlockFinisher(z.size).finish(dz, range_cor=drc, data_type="complex32")
    assert enc.dtype == torch.complex32 and enc.is_cuda
    assert np.array_equal(dz.cpu().numpy().view(np.uint32), zo.view(np.uint32))
    assert np.array_equal(enc.view(torch.int32).cpu().numpy().view(np.uint16),
                          enc_o.view(np.uint16))
    check_stats(st, ex)


@pytest.mark.parametrize("data_type", ["complex64", "complex32", "complex64_zero_mantissa"])
def test_background_writer_accumulates_revisited_blocks(torch_gpu, data_type):
    """Two passes over the same block (mixed-mode focusing) add up; other
    blocks are written as is; stats merge over the visited blocks."""
    from isce3_amd.rslc import BackgroundWriter, complex32
    rng = np.random.default_rng(12)
    shape = (96, 160)
    rc_full = np.exp(1j * rng.uniform(0, 6, shape[1])).astype(np.complex64)
    ds = FakeDataset(shape, complex32 if data_type == "complex32" else np.complex64)
    w = BackgroundWriter(rc_full, ds, data_type, mantissa_nbits=12)
    b1, b2 = np.s_[0:48, 0:160], np.s_[48:96, 0:160]
    z1, z1b, z2 = random_block(rng, 48, 160), random_block(rng, 48, 160), random_block(rng, 48, 160)
    w.queue_write(z1.copy(), b1)
    w.queue_write(z2.copy(), b2)
    w.notify_finished()
    w.write(z1b.copy(), b1)                  # second pass, in the caller's thread
    # oracle: first pass, readback of what was stored, second pass
    def store(z, prior=None):
        zo, enc, st = ro.finish_block(z, rc_full, prior, data_type, 12)
        return enc, st
    e1, _ = store(z1)
    prior = (e1["r"].astype(np.float32) + 1j * e1["i"].astype(np.float32)).astype(np.complex64) \
        if data_type == "complex32" else e1
    e1b, st1b = store(z1b, prior)
    e2, st2 = store(z2)
    got1, got2 = ds.data[b1], ds.data[b2]
    view = np.uint16 if data_type == "complex32" else np.uint32
    assert np.array_equal(got1.view(view), e1b.view(view))
    assert np.array_equal(got2.view(view), e2.view(view))
    total = w.stats
    assert total.n_valid == 2 * 48 * 160
    check_stats(w.block_stats(b1), st1b)
    check_stats(w.block_stats(b2), st2)


def test_full_size_block(torch_gpu):
    """A 4096 x 4096 block (the backprojection output size) with range
    correction and a prior block, complex32: bit-identical to the oracle."""
    torch = torch_gpu
    from isce3_amd.rslc import BlockFinisher
    rng = np.random.default_rng(99)
    n = 4096
    z = random_block(rng, n, n, 1e3)
    rc = np.exp(1j * rng.uniform(-np.pi, np.pi, n)).astype(np.complex64)
    prior = random_block(rng, n, n, 1e2)
    zo, enc_o, ex = ro.finish_block(z, rc, prior, "complex32")
    dz = torch.from_numpy(z).cuda()
    enc, st = BlockFinisher(n * n).finish(dz, range_cor=torch.from_numpy(rc).cuda(),
                                          prior=torch.from_numpy(prior).cuda(),
                                          data_type="complex32")
    assert np.array_equal(dz.cpu().numpy().view(np.uint32), zo.view(np.uint32))
    assert np.array_equal(enc.view(torch.int32).cpu().numpy().view(np.uint16),
                          enc_o.view(np.uint16))
    check_stats(st, ex)
