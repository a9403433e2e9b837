"""GPU parity of the HIP backprojector against the CPU oracle.

plain: reference semantics (FP64 delay/phase, float taps, complex<double> sum)
       vs the oracle in reference mode: complex relative error <= 1e-5 and
       RMS phase error <= 1e-5 rad over pixels within 40 dB of the peak.
split: FP64 per-pixel reference + FP32 deltas vs the all-FP64 oracle:
       RMS phase error <= 1e-3 rad (the target) and relative error <= 1e-3.
"""
import numpy as np
import pytest

from helpers import complex_rel_error, oracle_run, phase_errors, rms, scene_and_raw

pytestmark = pytest.mark.gpu

PLAIN_REL, PLAIN_PHS = 1e-5, 1e-5
SPLIT_REL, SPLIT_PHS = 1e-3, 1e-3


def run_gpu(config, precision, **kw):
    from isce3_amd.focus import backproject
    sc, raw = scene_and_raw(config)
    out = np.empty(sc.out_grid.shape, np.complex64)
    h = np.empty(sc.out_grid.shape, np.float32)
    stats = {}
    err = backproject(out, sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                      sc.kernel, sc.dry_tropo_model, height=h, precision=precision,
                      stats=stats, **kw)
    return out, h, err, stats


@pytest.mark.parametrize("config", ["tiny", "small", "pt"])
def test_plain_matches_reference_oracle(torch_gpu, config):
    print(f"[plain parity] config={config}", flush=True)
    out, h, err, stats = run_gpu(config, "plain")
    ref, href, _, rc = oracle_run(config, 0)
    assert err == (rc == 6)
    rel = complex_rel_error(out, ref)
    ph = phase_errors(out, ref)
    print(f"  rel={rel:.3e} rms_phase={rms(ph):.3e} max_phase={np.abs(ph).max():.3e}")
    assert rel <= PLAIN_REL
    assert rms(ph) <= PLAIN_PHS
    np.testing.assert_allclose(h, href, atol=1e-3)
    assert stats["updates"] > 0


@pytest.mark.parametrize("config", ["tiny", "small", "pt"])
def test_split_matches_fp64_oracle(torch_gpu, config):
    print(f"[split parity] config={config}", flush=True)
    out, h, err, stats = run_gpu(config, "split")
    ref, href, _, rc = oracle_run(config, 1)
    assert err == (rc == 6)
    rel = complex_rel_error(out, ref)
    ph = phase_errors(out, ref)
    print(f"  rel={rel:.3e} rms_phase={rms(ph):.3e} max_phase={np.abs(ph).max():.3e}")
    assert rel <= SPLIT_REL
    assert rms(ph) <= SPLIT_PHS
    np.testing.assert_allclose(h, href, atol=1e-3)


@pytest.mark.parametrize("tile_rows", [1, 2, 4, 8])
@pytest.mark.parametrize("precision", ["plain", "split"])
def test_tilings_agree(torch_gpu, tile_rows, precision):
    """Every wavefront tiling gives the plain result bit-for-bit (same per-pixel
    pulse order); split tilings agree within the split tolerance."""
    base, _, _, _ = run_gpu("tiny", precision)
    out, _, _, _ = run_gpu("tiny", precision, options={"tile_rows": tile_rows})
    if precision == "plain":
        np.testing.assert_array_equal(out, base)
    else:
        assert complex_rel_error(out, base) <= 1e-5


@pytest.mark.parametrize("config", ["tiny", "small", "pt"])
@pytest.mark.parametrize("tile_rows", [1, 2, 8])
def test_staged_windows_equal_global_gathers(torch_gpu, config, tile_rows):
    """The split sum with pulse windows staged in LDS gives the same result,
    bit for bit, as the same arithmetic with taps gathered from global memory.
    64-column tiles (tile_rows 1) span more range bins than one staged window
    holds and exercise the per-pulse fallback to global gathers."""
    print(f"[staged vs gather] config={config} tile_rows={tile_rows}", flush=True)
    staged, _, _, _ = run_gpu(config, "split", options={"tile_rows": tile_rows, "staged": 1})
    gathered, _, _, _ = run_gpu(config, "split", options={"tile_rows": tile_rows, "staged": 0,
                                                          "lane_rows": 1})
    np.testing.assert_array_equal(staged, gathered)


@pytest.mark.parametrize("config", ["tiny", "small"])
@pytest.mark.parametrize("tile_rows", [1, 2, 4, 8])
def test_two_lines_per_lane_match_fp64_oracle(torch_gpu, config, tile_rows):
    """Two adjacent output lines per lane sharing one 9-bin gather: within the
    split tolerance of the all-FP64 oracle and of the one-line kernel, for
    every tiling (odd line counts leave the second line of the last lane row
    outside the grid)."""
    print(f"[two lines] config={config} tile_rows={tile_rows}", flush=True)
    opts = {"tile_rows": tile_rows, "lane_rows": 2}
    out, h, err, _ = run_gpu(config, "split", options=opts)
    ref, href, _, rc = oracle_run(config, 1)
    assert err == (rc == 6)
    assert complex_rel_error(out, ref) <= SPLIT_REL
    assert rms(phase_errors(out, ref)) <= SPLIT_PHS
    np.testing.assert_allclose(h, href, atol=1e-3)
    one, _, _, _ = run_gpu(config, "split", options={"tile_rows": tile_rows, "lane_rows": 1})
    assert complex_rel_error(out, one) <= 1e-4
    assert np.array_equal(np.isnan(out), np.isnan(one))


@pytest.mark.parametrize("chunk", [16, 32, 64])
def test_fold_intervals_match_fp64_oracle(torch_gpu, chunk):
    """The phase steps of the split sum add the third and fourth differences
    once per fold interval (`chunk` pulses): every interval length the
    autotuner may pick stays within the split tolerance of the all-FP64
    oracle in both tilings, and the one-line kernels with staged and
    gathered taps (which take each pulse's phase one and two pulses ahead of
    accumulating it) still agree bit for bit."""
    print(f"[fold interval] chunk={chunk}", flush=True)
    ref, _, _, rc = oracle_run("small", 1)
    for lane_rows in (1, 2):
        out, _, err, _ = run_gpu("small", "split", options={"chunk": chunk, "lane_rows": lane_rows})
        assert err == (rc == 6)
        rel = complex_rel_error(out, ref)
        ph = rms(phase_errors(out, ref))
        print(f"  lane_rows={lane_rows} rel={rel:.3e} rms_phase={ph:.3e}", flush=True)
        assert rel <= SPLIT_REL
        assert ph <= SPLIT_PHS
    staged, _, _, _ = run_gpu("small", "split", options={"chunk": chunk, "staged": 1,
                                                          "tile_rows": 2})
    gathered, _, _, _ = run_gpu("small", "split", options={"chunk": chunk, "staged": 0,
                                                            "lane_rows": 1, "tile_rows": 2})
    np.testing.assert_array_equal(staged, gathered)


def test_fold_interval_limit(torch_gpu):
    """Fold intervals beyond 64 pulses (where the phase term the steps leave
    out would grow as the cube of the interval) are refused."""
    from isce3_amd import errors
    with pytest.raises(errors.InvalidArgument, match="chunk"):
        run_gpu("tiny", "split", options={"chunk": 65})
    with pytest.raises(errors.InvalidArgument, match="chunk"):
        run_gpu("tiny", "split", options={"chunk": -1})


def test_two_lines_per_lane_rejected_where_unsupported(torch_gpu):
    from isce3_amd import errors
    with pytest.raises(errors.InvalidArgument):
        run_gpu("tiny", "plain", options={"lane_rows": 2})
    with pytest.raises(errors.InvalidArgument):
        run_gpu("tiny", "split", options={"lane_rows": 2, "staged": 1})


def test_two_lines_per_lane_odd_kernel_width(torch_gpu):
    """The point-target scene uses Knab width 9 (the workflow's default
    halfwidth 4): the two-line kernel serves it against the FP64 oracle."""
    out, h, err, _ = run_gpu("pt", "split", options={"lane_rows": 2})
    ref, href, _, rc = oracle_run("pt", 1)
    assert err == (rc == 6)
    assert complex_rel_error(out, ref) <= SPLIT_REL
    assert rms(phase_errors(out, ref)) <= SPLIT_PHS


@pytest.mark.parametrize("entries", [4096, 8192])
def test_plain_kernel_tables_in_lds_and_global_memory(torch_gpu, entries):
    """The plain sum copies tables of up to 4096 entries to LDS and reads
    larger ones from global memory (its two instantiations): both against the
    oracle with the same table."""
    from isce3_amd.core import KnabKernel, TabulatedKernelF32
    from isce3_amd.focus import backproject
    import oracle
    sc, raw = scene_and_raw("tiny")
    kern = TabulatedKernelF32(KnabKernel(8.0, 0.8), entries)
    print(f"[plain table] entries={entries}", flush=True)
    out = np.empty(sc.out_grid.shape, np.complex64)
    err = backproject(out, sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                      kern, sc.dry_tropo_model, precision="plain")
    ref, _, _, rc = oracle.backproject(sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc,
                                       sc.ds, kern, sc.dry_tropo_model, mode=0)
    assert err == (rc == 6)
    rel, ph = complex_rel_error(out, ref), rms(phase_errors(out, ref))
    print(f"  rel={rel:.3e} rms_phase={ph:.3e}")
    assert rel <= PLAIN_REL and ph <= PLAIN_PHS


@pytest.mark.parametrize("width", [4, 7, 11, 16])
@pytest.mark.parametrize("precision,lane_rows", [("plain", 1), ("split", 1), ("split", 2)])
def test_kernel_widths_match_oracle(torch_gpu, width, precision, lane_rows):
    """Knab kernels of the widths the workflow can ask for (1 + 2 halfwidth and
    even widths; Backproject.cu:698-754 takes any kernel) against the oracle,
    in both precisions and both split tilings."""
    from isce3_amd.focus import backproject
    import oracle
    sc, raw = scene_and_raw("tiny", kernel_width=float(width))
    print(f"[width] {width} {precision} lane_rows={lane_rows}", flush=True)
    out = np.empty(sc.out_grid.shape, np.complex64)
    err = backproject(out, sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                      sc.kernel, sc.dry_tropo_model, precision=precision,
                      options={"lane_rows": lane_rows})
    mode = 0 if precision == "plain" else 1
    ref, _, _, rc = oracle.backproject(sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc,
                                       sc.ds, sc.kernel, sc.dry_tropo_model, mode=mode,
                                       kernel_bandwidth=0.8)
    assert err == (rc == 6)
    rel, ph = complex_rel_error(out, ref), rms(phase_errors(out, ref))
    print(f"  rel={rel:.3e} rms_phase={ph:.3e}")
    tol = (PLAIN_REL, PLAIN_PHS) if precision == "plain" else (SPLIT_REL, SPLIT_PHS)
    assert rel <= tol[0] and ph <= tol[1]
