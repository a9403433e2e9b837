"""Tile server on the GPU: each request's result is the same alone and in
the mix (bit for bit at equal launch options), and within the split-precision
tolerance of the FP64 oracle."""
import numpy as np
import pytest

from helpers import complex_rel_error, phase_errors, rms, scene_and_raw
from test_gpu_cases import gpu, orc

pytestmark = pytest.mark.gpu


def raw_sets(torch):
    from isce3_amd import serve
    sets, scenes = [], []
    for name in ("small", "tiny"):
        sc, raw = scene_and_raw(name)
        scenes.append((sc, raw))
        sets.append(serve.RawSet(torch.from_numpy(raw.copy()).cuda(), sc.in_geometry(), sc.orbit,
                                 sc.doppler, sc.dem, sc.fc, sc.ds, sc.kernel, sc.dry_tropo_model))
    return sets, scenes


def requests(scenes, n, seed):
    from isce3_amd import serve, sim
    reqs = []
    for i in range(n):
        sid = i % len(scenes)
        g = scenes[sid][0].out_grid
        (tile,) = sim.random_tiles(g, 1, 4, min(g.length, g.width), (0.5, 2.0), seed=seed + i)
        reqs.append(serve.Request(sid, tile, tag=i))
    return reqs


@pytest.mark.parametrize("precision,policy", [("split", "batch"), ("plain", "batch"),
                                              ("plain", "fcfs")])
def test_results_same_alone_and_in_the_mix(torch_gpu, precision, policy):
    from isce3_amd import serve
    sets, scenes = raw_sets(torch_gpu)
    opts = {"tile_rows": 2}
    srv = serve.Server(sets, precision=precision, policy=policy, workers=3, budget_bytes=8e9,
                       max_batch_pixels=1 << 18, keep_outputs=True, options=opts)
    reqs = srv.submit_many(requests(scenes, 18, seed=100))
    for r in reqs:
        r.wait(120)
    srv.close()
    if policy == "batch":
        assert max(b["requests"] for b in srv.batches) > 1, "no request was batched"
    else:
        assert all(b["requests"] == 1 for b in srv.batches)
    for r in reqs:
        sc, raw = scenes[r.set_id]
        alone, _, err = gpu(sc, raw, out_grid=r.grid, precision=precision, options=opts)
        np.testing.assert_array_equal(r.out.cpu().numpy(), alone)
        assert r.failed == err
        r.release()
    assert srv.in_use == srv.base_bytes


def test_served_split_within_tolerance_of_fp64_oracle(torch_gpu):
    from isce3_amd import serve
    sets, scenes = raw_sets(torch_gpu)
    srv = serve.Server(sets, precision="split", workers=2, budget_bytes=8e9,
                       max_batch_pixels=1 << 18, keep_outputs=True)
    reqs = requests(scenes, 6, seed=7)
    for r in reqs:
        srv.submit(r)
    for r in reqs:
        r.wait(120)
    srv.close()
    for r in reqs[:4]:
        sc, raw = scenes[r.set_id]
        ref, _, _ = orc(sc, raw, out_grid=r.grid, mode=1)
        z = r.out.cpu().numpy()
        if np.abs(ref).max() > 0:
            assert complex_rel_error(z, ref) <= 1e-3
            assert rms(phase_errors(z, ref)) <= 1e-3
