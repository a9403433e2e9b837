"""Batched focusing of several output blocks in one pass
(bp_backproject_blocks): every block's result equals a call of backproject on
that block alone, bit for bit, and the oracle's within the plain tolerance;
per-block failure flags are reported separately."""
import numpy as np
import pytest

from helpers import complex_rel_error, phase_errors, rms, scene_and_raw
from test_gpu_cases import gpu, orc

pytestmark = pytest.mark.gpu


def uneven_blocks(sc):
    """Blocks of differing shape, origin, line spacing and range spacing,
    including a single pixel, an empty block and ragged tile counts."""
    g = sc.out_grid
    specs = [(0, 0, 40, 33, 1.0, 1.0), (17.5, 100.25, 1, 1, 1.0, 1.0),
             (50, 3, 0, 7, 1.0, 1.0), (5, 60, 23, 129, 1.3, 0.7),
             (120, 10, 64, 64, 1.0, 1.0), (200, 150, 9, 200, 0.8, 0.37),
             (31, 31, 70, 3, 1.0, 2.0)]
    out = []
    for r0, c0, nl, nw, prf_f, dr_f in specs:
        b = g.offset_and_resize(r0, c0, nl, nw)
        b.prf = g.prf * prf_f
        b.range_pixel_spacing = g.range_pixel_spacing * dr_f
        out.append(b)
    return out


def run_blocks(sc, raw, grids, precision, **kw):
    from isce3_amd.focus import backproject_blocks
    outs = [np.full(b.shape, 5 + 5j, np.complex64) for b in grids]
    hs = [np.zeros(b.shape, np.float32) for b in grids]
    stats = {}
    flags = backproject_blocks(outs, [sc.out_geometry(b) for b in grids], raw, sc.in_geometry(),
                               sc.dem, sc.fc, sc.ds, sc.kernel, sc.dry_tropo_model,
                               heights=hs, precision=precision, stats=stats, **kw)
    return outs, hs, flags, stats


@pytest.mark.parametrize("precision,lane_rows", [("plain", 1), ("split", 1), ("split", 2)])
@pytest.mark.parametrize("tile_rows", [1, 2, 8])
def test_blocks_equal_single_calls(torch_gpu, precision, lane_rows, tile_rows):
    sc, raw = scene_and_raw("small")
    grids = uneven_blocks(sc)
    opts = {"tile_rows": tile_rows, "lane_rows": lane_rows}
    outs, hs, flags, stats = run_blocks(sc, raw, grids, precision, options=opts)
    total = 0
    for b, z, h, f in zip(grids, outs, hs, flags):
        st = {}
        z1, h1, f1 = gpu(sc, raw, out_grid=b, precision=precision, options=opts, stats=st)
        np.testing.assert_array_equal(z, z1)
        np.testing.assert_array_equal(h, h1)
        assert f == f1
        total += st.get("updates", 0)
    assert stats["updates"] == total


def test_blocks_match_oracle(torch_gpu):
    sc, raw = scene_and_raw("small")
    grids = uneven_blocks(sc)
    outs, hs, flags, _ = run_blocks(sc, raw, grids, "plain")
    for b, z, h in zip(grids, outs, hs):
        if b.length * b.width == 0:
            assert z.size == 0
            continue
        ref, href, _ = orc(sc, raw, out_grid=b)
        np.testing.assert_array_equal(np.isnan(z), np.isnan(ref))
        if np.abs(ref).max() > 0:
            assert complex_rel_error(z, ref) <= 1e-5
            assert rms(phase_errors(z, ref)) <= 1e-5
        np.testing.assert_allclose(h, href, atol=1e-3)


def test_split_blocks_match_fp64_oracle(torch_gpu):
    sc, raw = scene_and_raw("small")
    grids = uneven_blocks(sc)[3:6]
    outs, _, _, _ = run_blocks(sc, raw, grids, "split")
    for b, z in zip(grids, outs):
        ref, _, _ = orc(sc, raw, out_grid=b, mode=1)
        assert complex_rel_error(z, ref) <= 1e-3
        assert rms(phase_errors(z, ref)) <= 1e-3


def test_per_block_failure_flags(torch_gpu):
    """A look-angle interval that excludes the far block fails only that one."""
    sc, raw = scene_and_raw("tiny")
    from isce3_amd.geometry import rdr2geo_bracket
    g = sc.out_grid
    near = g.offset_and_resize(0, 0, 8, 10)
    far = g.offset_and_resize(0, g.width - 10, 8, 10)
    x_mid = rdr2geo_bracket(g.sensing_mid, g.slant_range(g.width * 0.5), sc.orbit, "right", 0.0,
                            g.wavelength, sc.dem)
    p, _ = sc.orbit.interpolate(g.sensing_mid)
    look = np.arccos(np.dot(-p / np.linalg.norm(p), (x_mid - p) / np.linalg.norm(x_mid - p)))
    r2g = {"look_min": 0.2, "look_max": float(look)}
    outs, hs, flags, stats = run_blocks(sc, raw, [near, far], "plain", rdr2geo_params=r2g)
    assert flags == [False, True]
    assert stats["any_failed"]
    assert np.isfinite(outs[0]).all() and np.isnan(outs[1]).all() and np.isnan(hs[1]).all()
