"""Resident scenes (bp_scene_*): the step replayed from a captured HIP graph
gives every block the result of bp_backproject_blocks bit for bit, on every
replay, with eager work and synchronisations between the replays, for
batches of different sizes and launch options, and rejects requests beyond
its reserved scratch."""
import numpy as np
import pytest

from helpers import scene_and_raw
from test_gpu_blocks import uneven_blocks

pytestmark = pytest.mark.gpu


def eager(sc, raw_dev, grids, precision, opts):
    import torch
    from isce3_amd.focus import backproject_blocks
    outs = [torch.full(tuple(b.shape), 5 + 5j, dtype=torch.complex64, device="cuda:0")
            for b in grids]
    hs = [torch.zeros(tuple(b.shape), dtype=torch.float32, device="cuda:0") for b in grids]
    st = {}
    flags = backproject_blocks(outs, [sc.out_geometry(b) for b in grids], raw_dev,
                               sc.in_geometry(), sc.dem, sc.fc, sc.ds, sc.kernel,
                               sc.dry_tropo_model, heights=hs, precision=precision,
                               options=opts, stats=st)
    return [o.cpu().numpy() for o in outs], [h.cpu().numpy() for h in hs], flags, st


def replay(scene, grids, opts, heights=True):
    import torch
    outs = [torch.full(tuple(b.shape), -1 - 1j, dtype=torch.complex64, device="cuda:0")
            for b in grids]
    hs = [torch.zeros(tuple(b.shape), dtype=torch.float32, device="cuda:0") for b in grids] \
        if heights else None
    st = {}
    flags = scene.focus(outs, grids, hs, options=opts, stats=st)
    return ([o.cpu().numpy() for o in outs], [h.cpu().numpy() for h in hs] if heights else None,
            flags, st)


@pytest.mark.parametrize("precision", ["plain", "split"])
def test_graph_replays_equal_eager_calls(torch_gpu, precision):
    from isce3_amd.focus import Scene
    torch = torch_gpu
    sc, raw = scene_and_raw("small")
    raw_dev = torch.from_numpy(raw.copy()).cuda()
    grids = uneven_blocks(sc)
    opts = {"tile_rows": 2}
    ref = eager(sc, raw_dev, grids, precision, opts)
    with Scene(raw_dev, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc, sc.ds,
               sc.kernel, sc.dry_tropo_model, precision=precision, max_pixels=1 << 16,
               max_blocks=16) as scene:
        for rep in range(3):
            print(f"[scene] precision={precision} replay={rep}", flush=True)
            outs, hs, flags, st = replay(scene, grids, opts)
            for z, z0, h, h0 in zip(outs, ref[0], hs, ref[1]):
                np.testing.assert_array_equal(z, z0)
                np.testing.assert_array_equal(h, h0)
            assert flags == ref[2]
            assert st["updates"] == ref[3]["updates"]
            # eager work and a synchronisation between replays
            eager(sc, raw_dev, grids[:2], precision, opts)
            torch.cuda.synchronize()
        # a smaller batch replays the same graph with new counts
        sub = [grids[3], grids[0]]
        ref_sub = eager(sc, raw_dev, sub, precision, opts)
        outs, _, flags, _ = replay(scene, sub, opts, heights=False)
        for z, z0 in zip(outs, ref_sub[0]):
            np.testing.assert_array_equal(z, z0)


def test_new_launch_options_capture_a_new_graph(torch_gpu):
    from isce3_amd.focus import Scene
    torch = torch_gpu
    sc, raw = scene_and_raw("tiny")
    raw_dev = torch.from_numpy(raw.copy()).cuda()
    grids = [sc.out_grid]
    with Scene(raw_dev, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc, sc.ds,
               sc.kernel, sc.dry_tropo_model, precision="split", max_pixels=1 << 14,
               max_blocks=4) as scene:
        for opts in ({"tile_rows": 1}, {"tile_rows": 8, "chunk": 16, "staged": 1},
                     {"tile_rows": 4, "blocks_per_cu": 1}, {"tile_rows": 2, "lane_rows": 2},
                     {"tile_rows": 1}):
            print(f"[scene] options={opts}", flush=True)
            ref = eager(sc, raw_dev, grids, "split", opts)
            outs, _, _, _ = replay(scene, grids, opts, heights=False)
            np.testing.assert_array_equal(outs[0], ref[0][0])


def test_scene_limits(torch_gpu):
    from isce3_amd import errors
    from isce3_amd.focus import Scene
    torch = torch_gpu
    sc, raw = scene_and_raw("tiny")
    raw_dev = torch.from_numpy(raw.copy()).cuda()
    g = sc.out_grid
    with Scene(raw_dev, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc, sc.ds,
               sc.kernel, sc.dry_tropo_model, precision="split", max_pixels=g.length * g.width - 1,
               max_blocks=1) as scene:
        out = torch.empty(tuple(g.shape), dtype=torch.complex64, device="cuda:0")
        with pytest.raises(errors.InvalidArgument):
            scene.focus([out], [g])
        small = g.offset_and_resize(0, 0, 4, 4)
        o1 = torch.empty((4, 4), dtype=torch.complex64, device="cuda:0")
        o2 = torch.empty((4, 4), dtype=torch.complex64, device="cuda:0")
        with pytest.raises(errors.InvalidArgument):
            scene.focus([o1, o2], [small, small])  # two blocks, one reserved
        assert scene.focus([o1], [small]) == [False]
