"""The range-compression oracle pinned to the reference's own known answers
(tests/cxx/isce3/focus/rangecomp.cpp, chirp.cpp,
tests/python/extensions/pybind/focus/rangecomp.py), to direct summation and to
the closed form of a compressed linear chirp."""
import numpy as np
import pytest

import rangecomp as rco


def test_next_fast_power():
    # brute force over 2^a 3^b 5^c
    smooth = sorted({2 ** a * 3 ** b * 5 ** c for a in range(16) for b in range(10)
                     for c in range(7) if 2 ** a * 3 ** b * 5 ** c < 2 ** 16})
    for n in [0, 1, 2, 7, 11, 13, 97, 1000, 1001, 4097, 16385, 20000, 33333]:
        want = 1 if n <= 1 else min(s for s in smooth if s >= n)
        assert rco.next_fast_power(n) == want


def test_chirp_matches_reference_checks():
    """tests/cxx/isce3/focus/chirp.cpp:9-62: sample count, amplitude, centre
    phase and the phase step between samples."""
    startfreq, endfreq, samples, spacing = 0.0, 20.0, 1001, 0.01
    amplitude, phi = 7.5, 0.5 * np.pi
    duration = (samples - 1) * spacing
    chirprate = (endfreq - startfreq) / duration
    centerfreq = startfreq + 0.5 * chirprate * duration
    c = rco.form_linear_chirp(chirprate, duration, 1.0 / spacing, centerfreq, amplitude, phi)
    assert len(c) == samples
    assert np.max(np.abs(amplitude - np.abs(c.astype(np.complex128)))) < 1e-5
    assert np.angle(c[len(c) // 2]) == pytest.approx(phi, abs=1e-6)
    i = np.arange(samples - 1)
    f = startfreq + chirprate * (i + 0.5) * spacing
    dphi = np.angle(c[1:].astype(np.complex128) * np.conj(c[:-1].astype(np.complex128)))
    assert np.max(np.abs(2 * np.pi * f * spacing - dphi)) < 1e-5
    with pytest.raises(ValueError):
        rco.form_linear_chirp(1.0, -1.0, 1.0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_convolve_modes_known_answers(precision):
    """tests/cxx/isce3/focus/rangecomp.cpp:137-193."""
    chirp = np.ones(5, np.complex64)
    x = np.ones(9, np.complex64)
    want = {rco.FULL: [1, 2, 3, 4, 5, 5, 5, 5, 5, 4, 3, 2, 1],
            rco.VALID: [5, 5, 5, 5, 5],
            rco.SAME: [3, 4, 5, 5, 5, 5, 5, 4, 3]}
    first = {rco.FULL: 4, rco.VALID: 0, rco.SAME: 2}
    for mode, w in want.items():
        y = rco.rangecompress(chirp, x, mode, precision)[0]
        assert np.max(np.abs(y - np.array(w))) < 1e-5
        assert rco.output_size(5, 9, mode) == len(w)
        assert rco.first_valid_sample(5, mode) == first[mode]


def test_unit_chirp_is_identity():
    """tests/python/extensions/pybind/focus/rangecomp.py:4-27."""
    h = np.ones(1, np.complex64)
    assert np.allclose(rco.rangecompress(h, np.ones((1, 1)))[0], [1])
    x = np.arange(10, dtype=np.complex64).reshape(10, 1)
    assert np.allclose(rco.rangecompress(h, x), x)


@pytest.mark.parametrize("mode", [rco.FULL, rco.VALID, rco.SAME])
def test_matches_direct_summation(mode):
    rng = np.random.default_rng(3)
    chirp = rco.form_linear_chirp(1e12, 20e-6, 24e6)          # 481 samples
    x = (rng.standard_normal((3, 1000)) + 1j * rng.standard_normal((3, 1000))).astype(np.complex64)
    y = rco.rangecompress(chirp, x, mode)
    d = rco.direct_correlation(chirp, x, mode)
    assert y.shape == d.shape
    assert np.max(np.abs(y - d)) / np.max(np.abs(d)) < 1e-12


def test_compressed_chirp_closed_form():
    """tests/cxx/isce3/focus/rangecomp.cpp:84-129: a replica of the chirp
    compresses to (T - |t|) sinc(K t (T - |t|)) near the peak."""
    chirprate, duration, fs = 100.0, 2.0, 2400.0
    chirp = rco.form_linear_chirp(chirprate, duration, fs)
    y = rco.rangecompress(chirp, chirp[None, :], rco.FULL, "fp32")[0] / fs
    n = len(y)
    T = len(chirp) / fs
    t = np.arange(n) / fs - 0.5 * (n - 1) / fs
    want = (T - np.abs(t)) * np.sinc(chirprate * t * (T - np.abs(t)))
    k = 200
    sl = slice(n // 2 - k, n // 2 + k + 1)
    assert np.max(np.abs(y[sl] - want[sl])) < 1e-6
