"""Server logic of isce3_amd.serve on the CPU, with a stand-in executor that
records what each batch would launch: memory admission never exceeds the
budget and is first-come-first-served, requests arriving while a batch runs
join a later batch, batches hold one raw set and respect the size limits,
errors reach every request of their batch, and the fcfs policy runs one
request at a time in arrival order."""
import threading
import time

import pytest

from isce3_amd import errors, serve
from isce3_amd.core import DateTime
from isce3_amd.product import RadarGridParameters


class FakeSet:
    def __init__(self, nbytes):
        self.nbytes = nbytes


class FakeExecutor:
    def __init__(self, server_ref, gate=None, fail_on=None, delay=0.0):
        self.server_ref = server_ref
        self.gate = gate
        self.fail_on = fail_on
        self.delay = delay
        self.calls = []
        self.lock = threading.Lock()
        self.started = []
        self.stopped = []

    def start_worker(self, index):
        self.started.append(index)

    def stop_worker(self, index):
        self.stopped.append(index)

    def run(self, index, raw, requests, precision, options):
        srv = self.server_ref[0]
        with self.lock:
            self.calls.append({"worker": index, "tags": [r.tag for r in requests],
                               "sets": {r.set_id for r in requests},
                               "pixels": sum(r.pixels for r in requests),
                               "in_use": srv.in_use, "options": dict(options)})
        if self.gate is not None:
            self.gate.wait(5)
        if self.delay:
            time.sleep(self.delay)
        if self.fail_on is not None and self.fail_on in [r.tag for r in requests]:
            raise RuntimeError("device error")
        for r in requests:
            r.out = ("tile", r.tag)
            r.failed = False
        return {"updates": sum(r.pixels for r in requests), "step_ms": 0.0}


SCRATCH = 4096 * serve.RECORD_BYTES_PER_PIXEL["split"]  # one worker, 4096-pixel batches


def grid(n, m=None):
    return RadarGridParameters(0.0, 0.24, 1520.0, 850e3, 6.25, "right", n, m or n,
                               DateTime(2025, 1, 1))


def make(sets=(0, 0), budget=10 ** 7, workers=1, ex_kw=None, **kw):
    ref = [None]
    ex = FakeExecutor(ref, **(ex_kw or {}))
    srv = serve.Server([FakeSet(b) for b in sets], budget_bytes=budget, workers=workers,
                       max_batch_pixels=kw.pop("max_batch_pixels", 4096), executor=ex, **kw)
    ref[0] = srv
    return srv, ex


def test_base_memory_over_budget_rejected():
    with pytest.raises(ValueError, match="exceed the budget"):
        make(sets=(10 ** 6,), budget=10 ** 6 + SCRATCH - 1)


def test_request_checks():
    srv, _ = make()
    with pytest.raises(ValueError, match="max_batch_pixels"):
        srv.submit(serve.Request(0, grid(100)))
    with pytest.raises(ValueError, match="no raw set"):
        srv.submit(serve.Request(5, grid(8)))
    srv.close()
    with pytest.raises(errors.RuntimeError_):
        srv.submit(serve.Request(0, grid(8)))


def test_continuous_batching_joins_waiting_requests():
    gate = threading.Event()
    srv, ex = make(ex_kw={"gate": gate})
    first = srv.submit(serve.Request(0, grid(8), tag="a"))
    while not ex.calls:
        time.sleep(0.001)
    later = [srv.submit(serve.Request(0, grid(8), tag=f"b{i}")) for i in range(5)]
    gate.set()
    for r in [first] + later:
        r.wait(5)
    srv.close()
    assert ex.calls[0]["tags"] == ["a"]
    assert ex.calls[1]["tags"] == [f"b{i}" for i in range(5)]
    assert all(r.batch_size == 5 for r in later)
    assert ex.stopped == [0]


def test_batches_hold_one_set_and_respect_limits():
    gate = threading.Event()
    srv, ex = make(sets=(0, 0, 0), ex_kw={"gate": gate}, max_batch_requests=3)
    block = srv.submit(serve.Request(0, grid(8), tag="block"))
    while not ex.calls:
        time.sleep(0.001)
    reqs = [srv.submit(serve.Request(i % 3, grid(20), tag=f"r{i}")) for i in range(12)]
    gate.set()
    for r in [block] + reqs:
        r.wait(5)
    srv.close()
    for c in ex.calls:
        assert len(c["sets"]) == 1
        assert c["pixels"] <= 4096
        assert len(c["tags"]) <= 3
    # the set of the oldest waiting request goes first
    assert ex.calls[1]["tags"][0] == "r0"
    assert sorted(t for c in ex.calls[1:] for t in c["tags"]) == sorted(r.tag for r in reqs)


def test_pixel_limit_splits_batches():
    gate = threading.Event()
    srv, ex = make(ex_kw={"gate": gate})
    block = srv.submit(serve.Request(0, grid(8), tag="block"))
    while not ex.calls:
        time.sleep(0.001)
    reqs = [srv.submit(serve.Request(0, grid(40), tag=f"r{i}")) for i in range(5)]  # 1600 px
    gate.set()
    for r in [block] + reqs:
        r.wait(5)
    srv.close()
    assert [len(c["tags"]) for c in ex.calls[1:]] == [2, 2, 1]


def test_admission_respects_budget_first_come_first_served():
    """Outputs are kept until released: a big request at the head of the queue
    waits for memory and later small requests do not overtake it."""
    srv, ex = make(budget=SCRATCH + 40_000, keep_outputs=True, max_batch_pixels=4096)
    a = srv.submit(serve.Request(0, grid(60), tag="a"))       # 28,800 B
    a.wait(5)
    big = srv.submit(serve.Request(0, grid(50), tag="big"))   # 20,000 B: does not fit yet
    small = srv.submit(serve.Request(0, grid(5), tag="small"))  # would fit, must wait
    time.sleep(0.05)
    assert not big.done() and not small.done()
    assert srv.in_use <= srv.budget_bytes
    a.release()
    big.wait(5)
    small.wait(5)
    srv.close()
    assert [c["tags"] for c in ex.calls][1][0] == "big"
    assert srv.peak_in_use <= srv.budget_bytes
    assert all(c["in_use"] <= srv.budget_bytes for c in ex.calls)


def test_outputs_released_when_not_kept():
    srv, _ = make(budget=SCRATCH + 40_000)
    reqs = [srv.submit(serve.Request(0, grid(60), tag=i)) for i in range(4)]
    for r in reqs:
        r.wait(5)
    srv.close()
    assert srv.in_use == srv.base_bytes
    assert all(r.out is None for r in reqs)


def test_errors_reach_every_request_of_the_batch():
    gate = threading.Event()
    srv, ex = make(ex_kw={"gate": gate, "fail_on": "bad"})
    block = srv.submit(serve.Request(0, grid(8), tag="block"))
    while not ex.calls:
        time.sleep(0.001)
    r1 = srv.submit(serve.Request(0, grid(8), tag="bad"))
    r2 = srv.submit(serve.Request(0, grid(8), tag="ok"))
    gate.set()
    block.wait(5)
    for r in (r1, r2):
        with pytest.raises(RuntimeError, match="device error"):
            r.wait(5)
    r3 = srv.submit(serve.Request(0, grid(8), tag="after"))
    r3.wait(5)
    srv.close()
    assert r3.error is None and r3.failed is False
    assert srv.in_use == srv.base_bytes


def test_fcfs_policy_runs_one_request_at_a_time_in_order():
    srv, ex = make(sets=(0, 0), workers=4, policy="fcfs", ex_kw={"delay": 0.002})
    reqs = [srv.submit(serve.Request(i % 2, grid(8), tag=i)) for i in range(8)]
    for r in reqs:
        r.wait(5)
    srv.close()
    assert srv.workers == 1
    assert [c["tags"] for c in ex.calls] == [[i] for i in range(8)]


def test_several_workers_overlap():
    srv, ex = make(sets=(0, 0, 0), workers=3, ex_kw={"delay": 0.05})
    reqs = [srv.submit(serve.Request(i, grid(8), tag=i)) for i in range(3)]
    for r in reqs:
        r.wait(5)
    srv.close()
    starts = sorted(b["start"] for b in srv.batches)
    ends = sorted(b["end"] for b in srv.batches)
    assert starts[-1] < ends[0], "the three batches should run at the same time"
    assert {c["worker"] for c in ex.calls} == {0, 1, 2}


def test_options_callable_sees_the_batch():
    seen = []

    def opts(set_id, batch):
        seen.append((set_id, len(batch)))
        return {"tile_rows": 2}

    srv, ex = make(options=opts)
    srv.submit(serve.Request(1, grid(8))).wait(5)
    srv.close()
    assert seen == [(1, 1)] and ex.calls[0]["options"] == {"tile_rows": 2}


def test_poisson_load_and_percentiles():
    srv, _ = make(sets=(0,), workers=2)
    reqs = [serve.Request(0, grid(8), tag=i) for i in range(50)]
    res = serve.poisson_load(srv, reqs, rate=2000.0, seed=3)
    srv.close()
    assert res["requests"] == 50 and res["p50_ms"] <= res["p99_ms"]
    assert all(r.latency >= 0 for r in reqs)
    assert serve.percentile([1, 2, 3, 4], 50) == 2 and serve.percentile([1, 2, 3, 4], 99) == 4
    assert serve.percentile(list(range(1, 101)), 99) == 99


def test_submit_many_lands_in_one_batch():
    srv, ex = make()
    reqs = srv.submit_many([serve.Request(0, grid(8), tag=i) for i in range(6)])
    for r in reqs:
        r.wait(5)
    srv.close()
    assert [c["tags"] for c in ex.calls] == [list(range(6))]
