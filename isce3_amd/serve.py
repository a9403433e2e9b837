"""Tile-focusing server: raw pulse sets resident in HBM, requests admitted by
device memory, continuous batching onto concurrent HIP streams.

A request asks for one output tile (a radar grid) focused from one resident
raw set.  Requests move through three stages:

  arrived   queued in arrival order (``submit``)
  admitted  their output memory fits the HBM budget; admission is strictly
            first-come-first-served, so a large request at the head is not
            overtaken by later small ones
  running   taken by an idle worker as part of a batch

Each worker owns a HIP stream and a native context.  When it becomes idle it
forms a batch from the admitted requests: all admitted requests of the raw
set whose oldest request has waited longest, in arrival order, up to
``max_batch_pixels`` and ``max_batch_requests``.  The batch is focused by one
``bp_backproject_blocks`` call (one geometry launch + one persistent
coherent-sum launch over every tile of the batch), so requests that arrive
while other batches run join the next batch, and batches of different
workers overlap on the device.  The ``fcfs`` policy (one worker, one request
per batch) is the first-come-first-served baseline.

Memory accounting: the resident raw sets and each worker's per-pixel scratch
(sized for ``max_batch_pixels``, one copy per raw set when the executor keeps
a resident scene per set) are charged once at construction; every admitted
request is charged its output bytes until it is released (at once when
``keep_outputs`` is false, else by ``Request.release``).
"""
from __future__ import annotations

import collections
import ctypes as C
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

from . import _native, errors

OUT_BYTES_PER_PIXEL = 8                        # complex64 output
# device scratch per pixel of a context: geometry (+ split) records; equal to
# bp_scratch_bytes_per_pixel (checked by the tests against the library)
RECORD_BYTES_PER_PIXEL = {"plain": 48, "split": 48 + 64}


@dataclass
class RawSet:
    """One range-compressed pulse set resident on the device and everything
    a tile request needs besides its output grid."""
    signal: object                 # device complex64 tensor [pulses x bins]
    in_geometry: object            # RadarGeometry of the pulses
    out_orbit: object              # Orbit of the output grids
    out_doppler: object            # LUT2d of the output grids (zero Doppler: LUT2d())
    dem: object                    # DEMInterpolator
    fc: float
    ds: float
    kernel: object                 # TabulatedKernelF32
    dry_tropo_model: str = "tsx"
    rdr2geo_params: dict = field(default_factory=dict)
    geo2rdr_params: dict = field(default_factory=dict)

    @property
    def nbytes(self) -> int:
        s = self.signal
        if hasattr(s, "element_size"):
            return int(s.numel() * s.element_size())
        return int(s.nbytes)


class Request:
    """A tile-focusing request; ``wait`` returns it once focused."""

    def __init__(self, set_id: int, grid, arrival: Optional[float] = None, tag=None):
        self.set_id = int(set_id)
        self.grid = grid
        self.pixels = int(grid.length) * int(grid.width)
        self.arrival = time.perf_counter() if arrival is None else float(arrival)
        self.tag = tag
        self.admitted: Optional[float] = None
        self.started: Optional[float] = None
        self.finished: Optional[float] = None
        self.batch_id: Optional[int] = None
        self.batch_size: Optional[int] = None
        self.out = None
        self.failed: Optional[bool] = None
        self.error: Optional[BaseException] = None
        self._done = threading.Event()
        self._server: Optional["Server"] = None
        self._charged = 0

    @property
    def out_bytes(self) -> int:
        return self.pixels * OUT_BYTES_PER_PIXEL

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival

    def done(self) -> bool:
        return self._done.is_set()

    def wait(self, timeout: Optional[float] = None) -> "Request":
        if not self._done.wait(timeout):
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise self.error
        return self

    def release(self):
        """Give the output memory back to the server's budget."""
        self.out = None
        if self._server is not None:
            self._server._release(self)


class GpuExecutor:
    """Runs batches on the GPU: one HIP stream and one native context per
    worker.  With ``graphs`` (the default) each worker keeps one resident
    scene per raw set (focus.Scene: shared data uploaded once, the step
    replayed from a captured HIP graph); otherwise every batch is one
    bp_backproject_blocks call that uploads the shared data again."""

    def __init__(self, device: int = 0, graphs: bool = True):
        import torch  # deferred: the server logic itself needs no torch
        if not torch.cuda.is_available():
            raise errors.HipError("the tile server needs a ROCm GPU (none visible)")
        self.torch = torch
        self.device = int(device)
        self.graphs = bool(graphs)
        self._workers = {}
        self._scenes = {}
        self._limits = (1 << 24, 256)

    def scratch_copies(self, n_sets: int) -> int:
        """Per-pixel scratch buffers each worker holds (one per raw set when
        every set has its own scene)."""
        return n_sets if self.graphs else 1

    def configure(self, max_batch_pixels: int, max_batch_requests: int):
        self._limits = (int(max_batch_pixels), int(max_batch_requests))

    def start_worker(self, index: int):
        torch = self.torch
        with torch.cuda.device(self.device):
            stream = torch.cuda.Stream(device=self.device)  # default priority
        h = C.c_void_p()
        _native.check(_native.lib().bp_context_create(self.device, C.byref(h)))
        self._workers[index] = (stream, h)

    def stop_worker(self, index: int):
        for key in [k for k in self._scenes if k[0] == index]:
            self._scenes.pop(key).close()
        stream, h = self._workers.pop(index)
        _native.lib().bp_context_destroy(h)

    def _scene(self, index: int, raw: RawSet, precision: str):
        from .focus import Scene
        key = (index, id(raw), precision)
        sc = self._scenes.get(key)
        if sc is None:
            sc = Scene(raw.signal, raw.in_geometry, raw.out_orbit, raw.out_doppler, raw.dem,
                       raw.fc, raw.ds, raw.kernel, raw.dry_tropo_model, raw.rdr2geo_params,
                       raw.geo2rdr_params, precision=precision, max_pixels=self._limits[0],
                       max_blocks=self._limits[1])
            self._scenes[key] = sc
        return sc

    def run(self, index: int, raw: RawSet, requests: Sequence[Request], precision: str,
            options: dict) -> dict:
        from .focus import (_PRECISION, _options_struct, geometry_struct,
                            parse_dry_tropo_model, parse_geo2rdr_params, parse_rdr2geo_params)
        torch = self.torch
        stream, ctx = self._workers[index]
        with torch.cuda.device(self.device), torch.cuda.stream(stream):
            outs = [torch.empty((r.grid.length, r.grid.width), dtype=torch.complex64,
                                device=f"cuda:{self.device}") for r in requests]
        if self.graphs:
            st = {}
            failed = self._scene(index, raw, precision).focus(
                outs, [r.grid for r in requests], options=options, stream=stream, stats=st)
            for r, o, f in zip(requests, outs, failed):
                r.out = o
                r.failed = bool(f)
            return {"updates": int(st["updates"]), "step_ms": float(st["step_ms"])}
        blocks = (_native.OutBlock_t * len(requests))()
        for i, r in enumerate(requests):
            gs = _native.Grid_t(r.grid.sensing_start, r.grid.prf, r.grid.starting_range,
                                r.grid.range_pixel_spacing, r.grid.length, r.grid.width,
                                int(r.grid.lookside), 0)
            blocks[i].grid = gs
            blocks[i].out_dev = outs[i].data_ptr()
            blocks[i].height_dev = None
        oorb = raw.out_orbit._struct()
        odop = raw.out_doppler._struct()
        ig = geometry_struct(raw.in_geometry)
        dm = raw.dem._struct()
        kr = raw.kernel._struct()
        opt = _options_struct(options)
        st = _native.Stats_t()
        failed = C.c_int(0)
        rc = _native.lib().bp_backproject_blocks(
            ctx, C.byref(oorb), C.byref(odop), blocks, len(requests), C.byref(ig), C.byref(dm),
            float(raw.fc), float(raw.ds), C.byref(kr), parse_dry_tropo_model(raw.dry_tropo_model),
            C.byref(parse_rdr2geo_params(raw.rdr2geo_params)),
            C.byref(parse_geo2rdr_params(raw.geo2rdr_params)),
            C.c_void_p(raw.signal.data_ptr()), _PRECISION[precision], C.byref(opt),
            C.c_void_p(stream.cuda_stream), C.byref(failed), C.byref(st))
        _native.check(rc)
        for i, r in enumerate(requests):
            r.out = outs[i]
            r.failed = bool(blocks[i].failed)
        return {"updates": int(st.updates), "step_ms": float(st.step_ms)}


class Server:
    """Continuous-batching tile server over resident raw sets (module docstring)."""

    def __init__(self, sets: Sequence[RawSet], *, precision: str = "split",
                 budget_bytes: float = 200e9, workers: int = 3,
                 max_batch_pixels: int = 1 << 24, max_batch_requests: int = 256,
                 policy: str = "batch", options=None, keep_outputs: bool = False,
                 executor=None, device: int = 0):
        if precision not in RECORD_BYTES_PER_PIXEL:
            raise errors.InvalidArgument(f"precision must be 'plain' or 'split', got {precision!r}")
        if policy not in ("batch", "fcfs"):
            raise errors.InvalidArgument(f"policy must be 'batch' or 'fcfs', got {policy!r}")
        if workers < 1 or max_batch_pixels < 1 or max_batch_requests < 1:
            raise errors.InvalidArgument("workers and batch limits must be positive")
        self.sets = list(sets)
        self.precision = precision
        self.policy = policy
        self.workers = 1 if policy == "fcfs" else int(workers)
        self.max_batch_pixels = int(max_batch_pixels)
        self.max_batch_requests = 1 if policy == "fcfs" else int(max_batch_requests)
        self.keep_outputs = bool(keep_outputs)
        # options: a dict, or a callable (set_id, requests) -> dict
        self._options = options if options is not None else {}
        self.budget_bytes = int(budget_bytes)
        self.executor = executor if executor is not None else GpuExecutor(device)
        copies = self.executor.scratch_copies(len(self.sets)) \
            if hasattr(self.executor, "scratch_copies") else 1
        self.base_bytes = sum(s.nbytes for s in self.sets) + self.workers * copies * (
            self.max_batch_pixels * RECORD_BYTES_PER_PIXEL[precision])
        if self.base_bytes > self.budget_bytes:
            raise errors.InvalidArgument(
                f"resident sets and worker scratch ({self.base_bytes} B) exceed the budget "
                f"({self.budget_bytes} B)")
        self.in_use = self.base_bytes
        self.peak_in_use = self.base_bytes
        self._cv = threading.Condition()
        self._arrived: collections.deque = collections.deque()
        self._admitted: List[Request] = []
        self._running = 0
        self._closed = False
        self._batch_seq = 0
        self.batches: List[dict] = []
        if hasattr(self.executor, "configure"):
            self.executor.configure(self.max_batch_pixels, self.max_batch_requests)
        self._threads = []
        for i in range(self.workers):
            self.executor.start_worker(i)
            t = threading.Thread(target=self._worker, args=(i,), daemon=True,
                                 name=f"isce3_amd-serve-{i}")
            t.start()
            self._threads.append(t)

    # ------------------------------------------------------------ client API
    def submit(self, req: Request) -> Request:
        return self.submit_many([req])[0]

    def submit_many(self, reqs: Sequence[Request]) -> List[Request]:
        """Submit requests together: they are queued in one step, so an idle
        worker sees all of them when it forms its next batch."""
        reqs = list(reqs)
        for req in reqs:
            self._check(req)
        with self._cv:
            if self._closed:
                raise errors.RuntimeError_("server is closed")
            for req in reqs:
                req._server = self
                self._arrived.append(req)
            self._admit_locked()
            self._cv.notify_all()
        return reqs

    def _check(self, req: Request):
        if not 0 <= req.set_id < len(self.sets):
            raise errors.InvalidArgument(f"no raw set {req.set_id}")
        if req.pixels > self.max_batch_pixels:
            raise errors.InvalidArgument(
                f"request of {req.pixels} pixels exceeds max_batch_pixels={self.max_batch_pixels}")
        if self.base_bytes + req.out_bytes > self.budget_bytes:
            raise errors.InvalidArgument("request output can never fit the memory budget")

    def close(self, wait: bool = True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            for t in self._threads:
                t.join()
            for i in range(self.workers):
                self.executor.stop_worker(i)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def pending(self) -> int:
        with self._cv:
            return len(self._arrived) + len(self._admitted) + self._running

    # ------------------------------------------------------------- internals
    def _admit_locked(self):
        now = time.perf_counter()
        while self._arrived:
            req = self._arrived[0]
            if self.in_use + req.out_bytes > self.budget_bytes:
                break  # first come, first served: no overtaking
            self._arrived.popleft()
            self.in_use += req.out_bytes
            req._charged = req.out_bytes
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            req.admitted = now
            self._admitted.append(req)

    def _release(self, req: Request):
        with self._cv:
            if req._charged:
                self.in_use -= req._charged
                req._charged = 0
                self._admit_locked()
                self._cv.notify_all()

    def _form_batch_locked(self) -> List[Request]:
        if not self._admitted:
            return []
        head = min(self._admitted, key=lambda r: r.arrival)
        batch, px = [], 0
        for r in sorted(self._admitted, key=lambda r: r.arrival):
            if r.set_id != head.set_id:
                continue
            if batch and (px + r.pixels > self.max_batch_pixels
                          or len(batch) >= self.max_batch_requests):
                break
            batch.append(r)
            px += r.pixels
        taken = set(id(r) for r in batch)
        self._admitted = [r for r in self._admitted if id(r) not in taken]
        return batch

    def _options_for(self, set_id, batch):
        o = self._options
        return dict(o(set_id, batch)) if callable(o) else dict(o)

    def _worker(self, index: int):
        while True:
            with self._cv:
                while True:
                    batch = self._form_batch_locked()
                    if batch:
                        break
                    if self._closed and not self._arrived and not self._admitted:
                        return
                    self._cv.wait()
                self._running += 1
                self._batch_seq += 1
                bid = self._batch_seq
            t0 = time.perf_counter()
            for r in batch:
                r.started = t0
                r.batch_id = bid
                r.batch_size = len(batch)
            info, err = None, None
            try:
                info = self.executor.run(index, self.sets[batch[0].set_id], batch,
                                         self.precision, self._options_for(batch[0].set_id, batch))
            except BaseException as e:  # delivered to every request of the batch
                err = e
            t1 = time.perf_counter()
            rec = {"id": bid, "worker": index, "set": batch[0].set_id, "requests": len(batch),
                   "pixels": sum(r.pixels for r in batch), "start": t0, "end": t1,
                   "error": repr(err) if err else None}
            if info:
                rec.update(info)
            with self._cv:
                self.batches.append(rec)
                self._running -= 1
            for r in batch:
                r.finished = t1
                r.error = err
                if err is not None or not self.keep_outputs:
                    r.out = None
                    self._release(r)
                r._done.set()
            with self._cv:
                self._cv.notify_all()


def percentile(values: Sequence[float], q: float) -> float:
    """Nearest-rank percentile (q in [0, 100])."""
    v = sorted(values)
    if not v:
        return float("nan")
    k = max(0, min(len(v) - 1, int(-(-q * len(v) // 100)) - 1))
    return float(v[k])


def poisson_load(server: Server, requests: Sequence[Request], rate: float, seed: int = 0,
                 clock: Callable[[], float] = time.perf_counter,
                 sleep: Callable[[float], None] = time.sleep) -> dict:
    """Submit `requests` at Poisson arrival times of `rate` per second (open
    loop: arrivals do not wait for completions), wait for all, and summarise
    latency (arrival -> result) and throughput."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / rate, size=len(requests))
    t_start = clock()
    t = t_start
    for req, gap in zip(requests, gaps):
        t += gap
        delay = t - clock()
        if delay > 0:
            sleep(delay)
        req.arrival = t
        server.submit(req)
    for req in requests:
        req.wait()
    t_end = max(r.finished for r in requests)
    lat = [r.latency for r in requests]
    return {"rate": rate, "requests": len(requests),
            "p50_ms": percentile(lat, 50) * 1e3, "p99_ms": percentile(lat, 99) * 1e3,
            "mean_ms": float(sum(lat) / len(lat)) * 1e3,
            "makespan_s": t_end - (t_start + float(gaps[0])),
            "pixels": sum(r.pixels for r in requests)}
