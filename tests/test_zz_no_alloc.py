"""bp_backproject_noalloc through the C ABI.

The entry point places every device buffer of a call in the library's fixed
per-device scratch array.  Checked here: its outputs (samples, heights, the
failure flag and the update count) equal bp_backproject's bit for bit on two
calls in a row, the second with a smaller raw set and output grid (so the
second reads a scratch array that still holds the first call's records), in
both precisions; a call made on a fresh context leaves the device's free
memory unchanged; a call whose records exceed the array is refused with
BP_ERR_LENGTH before any launch, and the next call still succeeds.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import scene_and_raw

pytestmark = pytest.mark.gpu


def _call(entry, ctx, torch, sc, d_raw, grid, precision):
    """One call of `entry` (bp_backproject or bp_backproject_noalloc) on the
    null stream; outputs start as sentinels so that unwritten pixels show."""
    from isce3_amd import _native
    from isce3_amd.focus import (geometry_struct, parse_dry_tropo_model, parse_geo2rdr_params,
                                 parse_rdr2geo_params)
    shape = tuple(grid.shape)
    out = torch.full(shape, complex(7.0, -7.0), dtype=torch.complex64, device="cuda:0")
    h = torch.full(shape, -999.0, dtype=torch.float32, device="cuda:0")
    ogeom, igeom = sc.out_geometry(grid), sc.in_geometry()
    og, ig = geometry_struct(ogeom), geometry_struct(igeom)
    dm, kr = sc.dem._struct(), sc.kernel._struct()
    r2g, g2r = parse_rdr2geo_params({}), parse_geo2rdr_params({})
    opt = _native.Options_t()
    st = _native.Stats_t()
    failed = C.c_int(0)
    prec = _native.PRECISION_SPLIT if precision == "split" else _native.PRECISION_PLAIN
    torch.cuda.synchronize()
    print(f"[noalloc] {entry} {precision} raw {tuple(d_raw.shape)} -> out {shape}", flush=True)
    rc = getattr(_native.lib(), entry)(
        ctx, C.byref(og), C.byref(ig), C.byref(dm), float(sc.fc), float(sc.ds), C.byref(kr),
        parse_dry_tropo_model(sc.dry_tropo_model), C.byref(r2g), C.byref(g2r),
        C.c_void_p(d_raw.data_ptr()), C.c_void_p(out.data_ptr()), C.c_void_p(h.data_ptr()),
        prec, C.byref(opt), None, C.byref(failed), C.byref(st))
    torch.cuda.synchronize()
    del ogeom, igeom
    return rc, out.cpu().numpy(), h.cpu().numpy(), failed.value, int(st.updates)


@pytest.fixture(scope="module")
def ctx(torch_gpu):
    """A context of its own, so that its grow-only arena starts empty."""
    from isce3_amd import _native
    torch_gpu.cuda.set_device(0)
    h = C.c_void_p()
    _native.check(_native.lib().bp_context_create(0, C.byref(h)))
    yield h
    _native.lib().bp_context_destroy(h)


@pytest.fixture(scope="module")
def raws(torch_gpu):
    out = {}
    for name in ("small", "tiny"):
        sc, raw = scene_and_raw(name)
        out[name] = (sc, torch_gpu.from_numpy(raw.copy()).to("cuda:0"))
    return out


@pytest.mark.parametrize("precision", ["split", "plain"])
def test_same_results_as_bp_backproject(torch_gpu, ctx, raws, precision):
    calls = []
    # larger raw set and grid first, then a smaller one: the second call's
    # records overwrite only part of what the first left in the array
    for name in ("small", "tiny"):
        sc, d_raw = raws[name]
        calls.append((name, _call("bp_backproject_noalloc", ctx, torch_gpu, sc, d_raw,
                                  sc.out_grid, precision)))
    for name, got in calls:
        sc, d_raw = raws[name]
        want = _call("bp_backproject", ctx, torch_gpu, sc, d_raw, sc.out_grid, precision)
        assert got[0] == want[0] == 0, (name, got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1], err_msg=f"{name} samples")
        np.testing.assert_array_equal(got[2], want[2], err_msg=f"{name} heights")
        assert got[3] == want[3] and got[4] == want[4] > 0
        assert np.isfinite(got[1]).all()


def test_no_device_memory_taken(torch_gpu, raws):
    """On a fresh context (empty arena) the call changes no free device memory."""
    from isce3_amd import _native
    h = C.c_void_p()
    _native.check(_native.lib().bp_context_create(0, C.byref(h)))
    try:
        sc, d_raw = raws["small"]
        assert _call("bp_backproject_noalloc", h, torch_gpu, sc, d_raw, sc.out_grid,
                     "split")[0] == 0
        free0 = torch_gpu.cuda.mem_get_info(0)[0]
        for precision in ("split", "plain"):
            rc = _call("bp_backproject_noalloc", h, torch_gpu, sc, d_raw, sc.out_grid,
                       precision)[0]
            assert rc == 0
        free1 = torch_gpu.cuda.mem_get_info(0)[0]
        # the outputs and sentinels come from torch's caching allocator, which
        # already holds blocks of these sizes from the warm-up call
        assert free1 == free0, (free0, free1)
    finally:
        _native.lib().bp_context_destroy(h)


def test_oversized_call_refused_then_next_call_works(torch_gpu, ctx, raws):
    from isce3_amd import _native
    from isce3_amd.product import RadarGridParameters
    sc, d_raw = raws["small"]
    g = sc.out_grid
    # 2048 x 2048 split records are 470 MB, beyond the 128 MiB array
    big = RadarGridParameters(g.sensing_start, g.wavelength, g.prf, g.starting_range,
                              g.range_pixel_spacing, g.lookside, 2048, 2048, g.ref_epoch)
    rc = _call("bp_backproject_noalloc", ctx, torch_gpu, sc, d_raw, big, "split")[0]
    assert rc == _native.BP_ERR_LENGTH
    assert "scratch" in _native.last_error()
    got = _call("bp_backproject_noalloc", ctx, torch_gpu, sc, d_raw, g, "split")
    want = _call("bp_backproject", ctx, torch_gpu, sc, d_raw, g, "split")
    assert got[0] == want[0] == 0
    np.testing.assert_array_equal(got[1], want[1])
