"""Range compression on the GPU against the oracle (FP64 restatement of
cxx/isce3/focus/RangeComp.cpp) and the reference's own known answers
(tests/cxx/isce3/focus/rangecomp.cpp, tests/python/extensions/pybind/focus/
rangecomp.py).  Tolerance: complex relative error <= 1e-5 against the FP64
oracle (FP32 FFTs of up to 2^19 points, or overlap-save blocks of 2^10 to
2^14 points)."""
import numpy as np
import pytest

import rangecomp as rco

pytestmark = pytest.mark.gpu

REL = 1e-5
MODES = {"full": rco.FULL, "valid": rco.VALID, "same": rco.SAME}


def rel_err(z, ref):
    return float(np.linalg.norm(z - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("mode", ["full", "valid", "same"])
def test_convolve_modes_known_answers(torch_gpu, mode):
    from isce3_amd.focus import RangeComp
    chirp = np.ones(5, np.complex64)
    x = np.ones(9, np.complex64)
    want = {"full": [1, 2, 3, 4, 5, 5, 5, 5, 5, 4, 3, 2, 1], "valid": [5, 5, 5, 5, 5],
            "same": [3, 4, 5, 5, 5, 5, 5, 4, 3]}[mode]
    rc = RangeComp(chirp, 9, 1, RangeComp.Mode[mode.capitalize()])
    assert rc.output_size == len(want)
    assert rc.first_valid_sample == {"full": 4, "valid": 0, "same": 2}[mode]
    y = np.zeros(rc.output_size, np.complex64)
    rc.rangecompress(y, x)
    print(f"[rangecomp KAT] mode={mode} max err={np.abs(y - want).max():.2e}", flush=True)
    assert np.abs(y - np.array(want)).max() < 1e-5


def test_unit_chirp_is_identity(torch_gpu):
    from isce3_amd.focus import RangeComp
    rc = RangeComp(np.ones(1, np.complex64), 1, maxbatch=10)
    assert (rc.chirp_size, rc.input_size, rc.maxbatch, rc.output_size) == (1, 1, 10, 1)
    assert rc.mode == RangeComp.Mode.Full and rc.fft_size >= 1
    x = np.ones(1, np.complex64)
    y = np.zeros_like(x)
    rc.rangecompress(y, x)
    assert np.allclose(y, x)
    x = np.arange(10, dtype=np.complex64).reshape(10, 1)
    y = np.zeros_like(x)
    rc.rangecompress(y, x)
    assert np.allclose(y, x)


def test_compressed_chirp_closed_form(torch_gpu):
    """A replica of the chirp compresses to (T - |t|) sinc(K t (T - |t|))
    (rangecomp.cpp:84-129; 9601-sample convolution, 2^14-point FFTs).  The
    reported fft_size is the reference's length, the smallest 2^a 3^b 5^c at or
    above 9601: 2^3 3^5 5 = 9720."""
    from isce3_amd.focus import RangeComp, form_linear_chirp
    chirprate, duration, fs = 100.0, 2.0, 2400.0
    chirp = form_linear_chirp(chirprate, duration, fs)
    rc = RangeComp(chirp, len(chirp))
    assert 2 * len(chirp) - 1 == 9601
    assert rc.fft_size == 9720 and rc.plan == ("fused", 16384, 1)
    y = np.zeros(rc.output_size, np.complex64)
    rc.rangecompress(y, chirp)
    y = y / fs
    n = len(y)
    T = len(chirp) / fs
    t = np.arange(n) / fs - 0.5 * (n - 1) / fs
    want = (T - np.abs(t)) * np.sinc(chirprate * t * (T - np.abs(t)))
    sl = slice(n // 2 - 200, n // 2 + 201)
    err = np.abs(y[sl] - want[sl]).max()
    print(f"[rangecomp chirp] max err={err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("mode", ["full", "valid", "same"])
@pytest.mark.parametrize("nin,batch", [(1000, 3), (3000, 5), (12000, 2), (700, 1)])
def test_random_pulses_match_fp64_oracle(torch_gpu, mode, nin, batch):
    from isce3_amd.focus import RangeComp, form_linear_chirp
    rng = np.random.default_rng(nin + batch)
    chirp = form_linear_chirp(1e12, 20e-6, 24e6)                # 481 samples
    x = (rng.standard_normal((batch, nin)) + 1j * rng.standard_normal((batch, nin))).astype(np.complex64)
    rc = RangeComp(chirp, nin, maxbatch=batch + 1, mode=RangeComp.Mode[mode.capitalize()])
    y = np.zeros((batch, rc.output_size), np.complex64)
    rc.rangecompress(y, x)
    ref = rco.rangecompress(chirp, x, MODES[mode], "fp64")
    e = rel_err(y, ref)
    print(f"[rangecomp random] mode={mode} nin={nin} batch={batch} fft={rc.fft_size} "
          f"rel={e:.2e}", flush=True)
    assert y.shape == ref.shape
    assert e <= REL
    assert rel_err(y, rco.direct_correlation(chirp, x, MODES[mode])) <= REL


def test_device_tensors_in_place_and_batch_limit(torch_gpu):
    from isce3_amd import errors
    from isce3_amd.focus import RangeComp, form_linear_chirp
    torch = torch_gpu
    rng = np.random.default_rng(5)
    chirp = form_linear_chirp(2e12, 10e-6, 48e6)
    x = (rng.standard_normal((4, 2000)) + 1j * rng.standard_normal((4, 2000))).astype(np.complex64)
    rc = RangeComp(chirp, 2000, maxbatch=4, mode=RangeComp.Mode.Same)
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.zeros((4, rc.output_size), dtype=torch.complex64, device="cuda:0")
    rc.rangecompress(d_out, d_in)
    ref = rco.rangecompress(chirp, x, rco.SAME, "fp64")
    assert rel_err(d_out.cpu().numpy(), ref) <= REL
    with pytest.raises(errors.LengthError):
        rc.rangecompress(np.zeros((5, rc.output_size), np.complex64),
                         np.zeros((5, 2000), np.complex64))
    with pytest.raises(errors.LengthError):
        rc.rangecompress(np.zeros((4, rc.output_size + 1), np.complex64), x)


@pytest.mark.parametrize("mode", ["full", "valid", "same"])
@pytest.mark.parametrize("nin,batch,duration", [(20000, 3, 40e-6), (40000, 2, 40e-6),
                                                (100000, 1, 40e-6), (250000, 2, 40e-6),
                                                (500000, 1, 40e-6), (1200000, 1, 40e-6),
                                                (60000, 2, 341.3e-6)])
def test_long_pulses_overlap_save_match_fp64_oracle(torch_gpu, mode, nin, batch, duration):
    """Convolutions past 2^14 samples with chirps of at most 2^13 + 1 samples
    (overlap-save blocks; 961 samples and 8191 samples, next to the 2^13 + 1
    limit of 2^14-point blocks), including a pulse longer than any single
    transform."""
    from isce3_amd.focus import RangeComp, form_linear_chirp
    rng = np.random.default_rng(nin + 7 * batch)
    chirp = form_linear_chirp(0.6e12, duration, 24e6)
    assert len(chirp) <= 8193
    x = (rng.standard_normal((batch, nin)) + 1j * rng.standard_normal((batch, nin))).astype(np.complex64)
    rc = RangeComp(chirp, nin, maxbatch=batch, mode=RangeComp.Mode[mode.capitalize()])
    assert rc.fft_size >= nin + len(chirp) - 1 > 16384
    y = np.zeros((batch, rc.output_size), np.complex64)
    rc.rangecompress(y, x)
    ref = rco.rangecompress(chirp, x, MODES[mode], "fp64")
    e = rel_err(y, ref)
    print(f"[rangecomp long] mode={mode} nin={nin} batch={batch} fft={rc.fft_size} "
          f"rel={e:.2e}", flush=True)
    assert y.shape == ref.shape
    assert e <= REL


@pytest.mark.parametrize("m,nin,kind,length,nblk", [
    (65, 2000, "overlap-save", 1 << 10, 3), (481, 12000, "overlap-save", 1 << 11, 8),
    (961, 15000, "overlap-save", 1 << 12, 5), (2001, 15000, "overlap-save", 1 << 13, 3),
    (8193, 60000, "overlap-save", 1 << 14, 8), (481, 3000, "fused", 1 << 12, 1),
    (4801, 4801, "fused", 1 << 14, 1), (9601, 9601, "four-step", 1 << 15, 1)])
def test_every_transform_plan_matches_fp64_oracle(torch_gpu, m, nin, kind, length, nblk):
    """The compressor's cost model picks one fused transform per pulse,
    overlap-save blocks of 2^10 .. 2^14 points, or four steps; each plan (the
    block seams included) gives the FP64 oracle's "same" convolution, and the
    reported fft_size is the reference's length, the smallest 2^a 3^b 5^c
    that holds the full convolution (RangeComp.cpp:60)."""
    from isce3_amd.focus import RangeComp
    rng = np.random.default_rng(m + nin)
    chirp = np.exp(1j * np.pi * 3e-5 * (np.arange(m) - m / 2) ** 2).astype(np.complex64)
    x = (rng.standard_normal((3, nin)) + 1j * rng.standard_normal((3, nin))).astype(np.complex64)
    rc = RangeComp(chirp, nin, maxbatch=3, mode=RangeComp.Mode.Same)
    print(f"[rangecomp plan] m={m} nin={nin} plan={rc.plan}", flush=True)
    assert rc.plan == (kind, length, nblk)
    assert rc.fft_size == rco.next_fast_power(m + nin - 1)
    y = np.zeros((3, rc.output_size), np.complex64)
    rc.rangecompress(y, x)
    ref = rco.rangecompress(chirp, x, rco.SAME, "fp64")
    e = rel_err(y, ref)
    print(f"[rangecomp plan] rel={e:.2e}", flush=True)
    assert e <= REL


@pytest.mark.parametrize("mode", ["full", "valid", "same"])
@pytest.mark.parametrize("nin,batch,m", [(20000, 3, 9001), (40000, 2, 12001), (150000, 1, 20001)])
def test_long_chirps_four_step_match_fp64_oracle(torch_gpu, mode, nin, batch, m):
    """Chirps longer than 2^13 + 1 samples: FFT lengths 2^15 .. 2^18 through
    the four-step path."""
    from isce3_amd.focus import RangeComp
    rng = np.random.default_rng(nin + m)
    chirp = np.exp(1j * np.pi * 1e-5 * (np.arange(m) - m / 2) ** 2).astype(np.complex64)
    x = (rng.standard_normal((batch, nin)) + 1j * rng.standard_normal((batch, nin))).astype(np.complex64)
    rc = RangeComp(chirp, nin, maxbatch=batch, mode=RangeComp.Mode[mode.capitalize()])
    y = np.zeros((batch, rc.output_size), np.complex64)
    rc.rangecompress(y, x)
    ref = rco.rangecompress(chirp, x, MODES[mode], "fp64")
    e = rel_err(y, ref)
    print(f"[rangecomp four-step] mode={mode} nin={nin} m={m} fft={rc.fft_size} rel={e:.2e}",
          flush=True)
    assert e <= REL


def test_long_pulses_more_than_one_work_batch(torch_gpu):
    """2^19-point FFTs of a long chirp: the 512 MiB work buffer holds 128
    pulses, so a batch of 130 runs as two sub-batches; pulses either side of
    the seam are checked."""
    from isce3_amd.focus import RangeComp
    torch = torch_gpu
    nin, batch, m = 300000, 130, 10001
    chirp = np.exp(1j * np.pi * 1e-5 * (np.arange(m) - m / 2) ** 2).astype(np.complex64)
    rc = RangeComp(chirp, nin, maxbatch=batch, mode=RangeComp.Mode.Same)
    assert rc.fft_size == rco.next_fast_power(nin + m - 1)
    assert rc.plan == ("four-step", 1 << 19, 1)
    g = torch.Generator(device="cuda:0").manual_seed(3)
    d_in = torch.randn((batch, nin), dtype=torch.complex64, device="cuda:0", generator=g)
    d_out = torch.empty((batch, rc.output_size), dtype=torch.complex64, device="cuda:0")
    rc.rangecompress(d_out, d_in)
    rows = [0, 1, 126, 127, 128, 129]
    x = d_in[rows].cpu().numpy()
    y = d_out[rows].cpu().numpy()
    ref = rco.rangecompress(chirp, x, rco.SAME, "fp64")
    for i, r in enumerate(rows):
        e = rel_err(y[i], ref[i])
        print(f"[rangecomp seam] pulse {r} rel={e:.2e}", flush=True)
        assert e <= REL
