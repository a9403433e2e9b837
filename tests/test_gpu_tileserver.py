"""The native tile server (bp_server_* / bp_serve_*): requests submitted from
several host threads, batched by the worker as they arrive, give every block
the result of bp_backproject_blocks bit for bit; closing the server finishes
every accepted request; oversized requests and unknown tickets are
rejected."""
import threading

import numpy as np
import pytest

from helpers import scene_and_raw
from test_gpu_blocks import uneven_blocks
from test_gpu_scene import eager

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["plain", "split"])
def test_concurrent_requests_match_eager_blocks(torch_gpu, precision):
    torch = torch_gpu
    from isce3_amd.focus import Scene, TileServer
    sc, raw = scene_and_raw("small")
    raw_dev = torch.from_numpy(raw.copy()).cuda()
    grids = uneven_blocks(sc)
    opts = {"tile_rows": 2}
    ref = eager(sc, raw_dev, grids, precision, opts)
    scene = Scene(raw_dev, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc, sc.ds,
                  sc.kernel, sc.dry_tropo_model, precision=precision, max_pixels=1 << 15,
                  max_blocks=4)
    outs = [torch.full(tuple(g.shape), 7 + 7j, dtype=torch.complex64, device="cuda:0")
            for g in grids]
    hs = [torch.zeros(tuple(g.shape), dtype=torch.float32, device="cuda:0") for g in grids]
    flags = [None] * len(grids)
    with TileServer(scene, options=opts) as server:
        def client(idx):
            tickets = [(i, server.submit(outs[i], grids[i], hs[i])) for i in idx]
            for i, t in tickets:
                flags[i] = server.wait(t)
        threads = [threading.Thread(target=client, args=(list(range(k, len(grids), 3)),))
                   for k in range(3)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
    scene.close()
    print(f"[tileserver] precision={precision} blocks={len(grids)}", flush=True)
    torch.cuda.synchronize()
    for o, h, z0, h0 in zip(outs, hs, ref[0], ref[1]):
        np.testing.assert_array_equal(o.cpu().numpy(), z0)
        np.testing.assert_array_equal(h.cpu().numpy(), h0)
    assert flags == ref[2]


def test_close_drains_and_rejections(torch_gpu):
    torch = torch_gpu
    from isce3_amd import errors
    from isce3_amd.focus import Scene, TileServer
    sc, raw = scene_and_raw("small")
    raw_dev = torch.from_numpy(raw.copy()).cuda()
    grids = uneven_blocks(sc)
    ref = eager(sc, raw_dev, grids, "split", None)
    scene = Scene(raw_dev, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc, sc.ds,
                  sc.kernel, sc.dry_tropo_model, precision="split", max_pixels=1 << 15,
                  max_blocks=8)
    server = TileServer(scene, max_pending=64)
    outs = [torch.zeros(tuple(g.shape), dtype=torch.complex64, device="cuda:0") for g in grids]
    tickets = [server.submit(o, g) for o, g in zip(outs, grids)]
    big = sc.out_grid
    with pytest.raises(errors.LengthError):
        server.submit(torch.zeros(tuple(big.shape), dtype=torch.complex64, device="cuda:0"), big)
    # poll until the first request completes, then close with the rest in flight
    while server.poll(tickets[0]) is None:
        pass
    with pytest.raises(errors.InvalidArgument):
        server.poll(tickets[0])          # reported once, then forgotten
    server.close()
    scene.close()
    torch.cuda.synchronize()
    for o, z0 in zip(outs, ref[0]):
        np.testing.assert_array_equal(o.cpu().numpy(), z0)


def test_bad_request_refused_alone(torch_gpu):
    """A request with PRF 0 is refused at submission; the requests around it
    are batched and focused as if it had never been sent."""
    torch = torch_gpu
    from isce3_amd import errors
    from isce3_amd.focus import Scene, TileServer
    sc, raw = scene_and_raw("small")
    raw_dev = torch.from_numpy(raw.copy()).cuda()
    grids = uneven_blocks(sc)
    ref = eager(sc, raw_dev, grids, "split", None)
    scene = Scene(raw_dev, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc, sc.ds,
                  sc.kernel, sc.dry_tropo_model, precision="split", max_pixels=1 << 15,
                  max_blocks=8)
    bad = grids[4].copy()
    bad.prf = 0.0
    outs = [torch.zeros(tuple(g.shape), dtype=torch.complex64, device="cuda:0") for g in grids]
    with TileServer(scene) as server:
        tickets = []
        for i, (o, g) in enumerate(zip(outs, grids)):
            tickets.append(server.submit(o, g))
            if i == 2:
                with pytest.raises(errors.InvalidArgument, match="PRF"):
                    server.submit(torch.zeros(tuple(bad.shape), dtype=torch.complex64,
                                              device="cuda:0"), bad)
        flags = [server.wait(t) for t in tickets]
    scene.close()
    torch.cuda.synchronize()
    assert flags == ref[2]
    for o, z0 in zip(outs, ref[0]):
        np.testing.assert_array_equal(o.cpu().numpy(), z0)


def test_two_waiters_on_one_ticket(torch_gpu):
    """A ticket has one consumer: of two threads waiting on it, one gets the
    result and the other an 'unknown ticket' error; neither hangs or crashes."""
    torch = torch_gpu
    from isce3_amd import errors
    from isce3_amd.focus import Scene, TileServer
    sc, raw = scene_and_raw("small")
    raw_dev = torch.from_numpy(raw.copy()).cuda()
    g = uneven_blocks(sc)[4]
    scene = Scene(raw_dev, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc, sc.ds,
                  sc.kernel, sc.dry_tropo_model, precision="split", max_pixels=1 << 15,
                  max_blocks=4)
    out = torch.zeros(tuple(g.shape), dtype=torch.complex64, device="cuda:0")
    results = []
    with TileServer(scene) as server:
        t = server.submit(out, g)

        def waiter():
            try:
                results.append(("ok", server.wait(t)))
            except errors.InvalidArgument as e:
                results.append(("unknown", str(e)))
        threads = [threading.Thread(target=waiter) for _ in range(2)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(60)
        assert not any(th.is_alive() for th in threads)
    scene.close()
    assert sorted(r[0] for r in results) == ["ok", "unknown"]
