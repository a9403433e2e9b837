"""The batched-tile and serving workloads at their full sizes.

Batched tiles: 64 output tiles (side log-uniform in [64, 1024] px, aspect
ratio uniform in [0.5, 2]) cut from one resident 8192 x 16384 raw set and
focused by one batched call (one geometry launch, one persistent coherent-sum
launch).  Every tile equals a single backproject call on that tile bit for
bit at the same launch options, in both precisions, and a pixel sample of
each tile meets the oracle's bar (plain: reference mode, 1e-5; split: FP64
mode, 1e-3).

Serving: three resident 8192 x 16384 raw sets, square tile requests (side
log-uniform in [32, 2048] px, random set) arriving as a Poisson stream and
continuously batched; each served result equals the same request focused
alone, and a pixel sample meets the split bar against the FP64 oracle.
"""
import os

import numpy as np
import pytest

import oracle
from helpers import complex_rel_error, phase_errors, rms

pytestmark = pytest.mark.gpu

SWATH_SEEDS = (3, 4, 5)


def say(msg):
    print(f"[configs] {msg}", flush=True)


@pytest.fixture(scope="module")
def swaths(torch_gpu):
    """Three seeded 8192 x 16384 raw sets: (scene, pinned host raw, device raw)."""
    torch = torch_gpu
    from isce3_amd import sim
    out = []
    for seed in SWATH_SEEDS:
        say(f"simulating swath seed {seed}")
        sc = sim.make_scene("swath", seed=seed)
        host = torch.empty((sc.in_grid.length, sc.in_grid.width), dtype=torch.complex64,
                           pin_memory=True)
        sim.simulate(sc, out=host.numpy(), nthreads=min(16, os.cpu_count() or 1))
        out.append((sc, host, host.cuda()))
    torch.cuda.synchronize()
    return out


def single(torch, sc, d_raw, grid, precision, options):
    from isce3_amd.focus import backproject
    out = torch.empty(tuple(grid.shape), dtype=torch.complex64, device="cuda:0")
    st = {}
    failed = backproject(out, sc.out_geometry(grid), d_raw, sc.in_geometry(), sc.dem, sc.fc,
                         sc.ds, sc.kernel, sc.dry_tropo_model, precision=precision,
                         options=options, stats=st)
    return out.cpu().numpy(), failed, st["updates"]


def oracle_sample(sc, host_raw, grid, z, mode, n, seed):
    """Oracle values at n seeded pixels of `grid` and the matching GPU values."""
    rng = np.random.Generator(np.random.PCG64(seed))
    npx = grid.length * grid.width
    sub = np.sort(rng.choice(npx, size=min(n, npx), replace=False)).astype(np.int64)
    ref, _, _, rc = oracle.backproject(sc.out_geometry(grid), host_raw.numpy(), sc.in_geometry(),
                                       sc.dem, sc.fc, sc.ds, sc.kernel, sc.dry_tropo_model,
                                       mode=mode, subset=sub, kernel_bandwidth=0.8,
                                       nthreads=min(16, os.cpu_count() or 1))
    assert rc == 0
    return z.ravel()[sub], ref


@pytest.mark.parametrize("precision,options", [("split", None), ("split", {"tile_rows": 4}),
                                               ("plain", None)])
def test_64_uneven_tiles_one_launch(torch_gpu, swaths, precision, options):
    torch = torch_gpu
    from isce3_amd import sim
    from isce3_amd.focus import backproject_blocks
    sc, host, d_raw = swaths[0]
    tiles = sim.random_tiles(sc.out_grid, 64, 64, 1024, (0.5, 2.0), seed=64)
    px = sum(t.length * t.width for t in tiles)
    say(f"batch {precision} {options}: 64 tiles, {px} pixels")
    outs = [torch.empty(tuple(t.shape), dtype=torch.complex64, device="cuda:0") for t in tiles]
    st = {}
    flags = backproject_blocks(outs, [sc.out_geometry(t) for t in tiles], d_raw,
                               sc.in_geometry(), sc.dem, sc.fc, sc.ds, sc.kernel,
                               sc.dry_tropo_model, precision=precision, options=options, stats=st)
    say(f"batched: {st['updates']} updates, sum {st['sum_ms']:.1f} ms")
    assert not any(flags)
    total = 0
    for i, (t, o) in enumerate(zip(tiles, outs)):
        z1, f1, upd = single(torch, sc, d_raw, t, precision, options)
        np.testing.assert_array_equal(o.cpu().numpy(), z1, err_msg=f"tile {i}")
        assert not f1
        total += upd
    assert st["updates"] == total
    say("batched == single calls on all 64 tiles")
    mode, tol = (oracle.MODE_REFERENCE, 1e-5) if precision == "plain" else (oracle.MODE_FP64, 1e-3)
    zs, refs = [], []
    for i in (0, 7, 21, 42, 63):
        z, ref = oracle_sample(sc, host, tiles[i], outs[i].cpu().numpy(), mode, 48, seed=i)
        zs.append(z)
        refs.append(ref)
    z, ref = np.concatenate(zs), np.concatenate(refs)
    assert np.isfinite(z).all()
    assert complex_rel_error(z, ref) <= tol
    assert rms(phase_errors(z, ref)) <= tol
    say(f"oracle sample rel {complex_rel_error(z, ref):.2e}")


def test_served_poisson_stream(torch_gpu, swaths):
    torch = torch_gpu
    from isce3_amd import serve, sim
    sets = [serve.RawSet(d, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc, sc.ds,
                         sc.kernel, sc.dry_tropo_model) for sc, _, d in swaths]
    rng = np.random.Generator(np.random.PCG64(40))
    reqs = []
    for i in range(12):
        sid = int(rng.integers(0, len(sets)))
        (tile,) = sim.random_tiles(swaths[sid][0].out_grid, 1, 32, 2048, (1.0, 1.0),
                                   seed=40 * 100003 + i)
        reqs.append(serve.Request(sid, tile, tag=i))
    say(f"serving 12 requests, {sum(r.pixels for r in reqs)} pixels")
    with serve.Server(sets, precision="split", workers=3, budget_bytes=200e9,
                      keep_outputs=True) as srv:
        # arrivals faster than service, so that requests queue and batch
        res = serve.poisson_load(srv, reqs, 200.0, seed=41)
        batches = list(srv.batches)
    say(f"served: p50 {res['p50_ms']:.1f} ms, p99 {res['p99_ms']:.1f} ms, "
        f"{len(batches)} batches")
    for i, r in enumerate(reqs):
        sc, host, d_raw = swaths[r.set_id]
        z = r.out.cpu().numpy()
        alone, failed, _ = single(torch, sc, d_raw, r.grid, "split", None)
        np.testing.assert_array_equal(z, alone, err_msg=f"request {i}")
        assert r.failed == failed and not failed
        if i < 4:
            zz, ref = oracle_sample(sc, host, r.grid, z, oracle.MODE_FP64, 64, seed=100 + i)
            assert complex_rel_error(zz, ref) <= 1e-3
            assert rms(phase_errors(zz, ref)) <= 1e-3
        r.release()
    say("served == alone on all 12 requests")
