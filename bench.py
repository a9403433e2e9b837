#!/usr/bin/env python3
"""Benchmark: pixel-pulse updates/s of the backprojector on one MI355X.

Headline workload (one "step" = one full backprojection of one block, inputs
resident in HBM): 16384 pulses x 16384 range bins complex64 (2 GiB) of a
seeded synthetic L-band stripmap scene (200 point targets + clutter, 1024^2
fractal DEM), focused to a 4096 x 4096 complex64 grid with a tabulated Knab
kernel (width 8, bandwidth 0.8, 2048 entries) at 6 m azimuth resolution.
`value` is the split-precision path (FP64 per-pixel reference + FP32
deltas); the plain all-FP64-inner-loop kernel is timed on the same inputs.
Accuracy of both is stated on a seeded pixel sample against the CPU oracle
(reference mode for plain, all-FP64 mode for split).  The CPU baseline is the
oracle in reference mode on a bounded sample of output rows.

Beside it, in the same JSON line:
  "batch"  64 output tiles (side log-uniform in [64, 1024] px, aspect ratio
           uniform in [0.5, 2], random positions in the illuminated swath)
           cut from one resident 8192 x 16384 raw set, focused by one batched
           call (one persistent sum launch, launch options from the
           autotuner) vs 64 sequential plain calls;
  "serve"  3 resident 8192 x 16384 raw sets, Poisson tile requests (square,
           side log-uniform in [32, 2048] px) at 20, 50 and 100 req/s,
           split precision with continuous batching vs the plain kernel
           served first-come-first-served; p50/p99 latency and updates/s.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_UPDATE = 101  # algorithmic flops per pixel-pulse update (FMA = 2)
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "sum_split_kernel_two_lines_pmc.md")
STAGE_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_rangecomp_rslc_presum.md")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def hbm_bytes_per_launch():
    """HBM bytes per launch of the split sum kernel on the headline workload,
    from the committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE
    scaled by a copy-kernel calibration); None when it is absent."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"HBM bytes per launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def stage_traffic(key):
    """HBM bytes per launch (per call for presum) of the range-compression,
    RSLC or presum kernel on its bench workload, from the committed counter
    summary; None when it is absent."""
    try:
        text = open(STAGE_TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"HBM bytes per (?:launch|call) \(" + key + r"\):\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def make_raw(torch, sim, sc, dev):
    raw_host = torch.empty((sc.in_grid.length, sc.in_grid.width), dtype=torch.complex64,
                           pin_memory=True)
    sim.simulate(sc, out=raw_host.numpy(), nthreads=min(16, os.cpu_count() or 1))
    raw = raw_host.to(dev)
    torch.cuda.synchronize()
    return raw_host, raw


# ------------------------------------------------------------- full block
def run_full(args, torch, np, dev, tuner):
    from isce3_amd import autotune, sim
    from isce3_amd.focus import backproject
    t0 = time.time()
    sc = sim.make_scene("full")
    raw_host, raw = make_raw(torch, sim, sc, dev)
    log(f"full scene: {len(sc.targets)} scatterers, {tuple(raw.shape)} simulated in "
        f"{time.time() - t0:.1f}s")
    out = torch.empty(sc.out_grid.shape, dtype=torch.complex64, device=dev)

    def step(precision, opts):
        st = {}
        backproject(out, sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                    sc.kernel, sc.dry_tropo_model, precision=precision, stats=st, options=opts)
        return st

    results = {}
    precisions = (args.only,) if args.only else ("plain", "split")
    for precision in precisions:
        if args.options is not None:
            opts = dict(args.options)
        elif args.no_tune:
            opts = None
        else:
            # launch options from the autotuner (untimed, part of the warmup)
            t0 = time.time()
            key = autotune.workload_key(device_name(torch), precision, 8,
                                        sc.out_grid.length * sc.out_grid.width,
                                        float(sc.out_grid.width))

            def device_ms(o, precision=precision):
                st = step(precision, o)
                return st["geometry_ms"] + st["sum_ms"]

            opts = tuner.tune(key, precision, device_ms, force=True)
            log(f"{precision}: autotuned {opts} in {time.time() - t0:.1f}s")
        for _ in range(args.warmup):
            step(precision, opts)
        torch.cuda.synchronize()
        stats = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            stats.append(step(precision, opts))
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        upd = stats[-1]["updates"]
        sum_ms = float(np.mean([s["sum_ms"] for s in stats]))
        geo_ms = float(np.mean([s["geometry_ms"] for s in stats]))
        results[precision] = dict(ms_per_step=dt * 1e3, updates=upd, sum_ms=sum_ms,
                                  geometry_ms=geo_ms, updates_per_s=upd / dt,
                                  options=opts or {}, out=out.cpu().numpy())
        log(f"{precision}: {dt * 1e3:.1f} ms/step, sum {sum_ms:.1f} ms, geometry {geo_ms:.1f} ms,"
            f" {upd / dt:.3e} updates/s")
    return sc, raw_host, results


def accuracy(args, np, sc, raw_np, results):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rng = np.random.Generator(np.random.PCG64(1234))
    npx = sc.out_grid.length * sc.out_grid.width
    subset = np.sort(rng.choice(npx, size=min(args.sample_pixels, npx), replace=False))
    nth = min(16, os.cpu_count() or 1)
    acc = {}
    for precision, mode in (("plain", oracle.MODE_REFERENCE), ("split", oracle.MODE_FP64)):
        if precision not in results or args.sample_pixels <= 0:
            continue
        ref, _, _, _ = oracle.backproject(sc.out_geometry(), raw_np, sc.in_geometry(), sc.dem,
                                          sc.fc, sc.ds, sc.kernel, sc.dry_tropo_model,
                                          mode=mode, subset=subset, nthreads=nth,
                                          kernel_bandwidth=0.8)
        z = results[precision]["out"].ravel()[subset]
        m = np.isfinite(ref) & np.isfinite(z)
        rel = float(np.linalg.norm(z[m] - ref[m]) / np.linalg.norm(ref[m]))
        a = np.abs(ref)
        strong = m & (a >= 1e-2 * a[m].max())
        ph = np.angle(z[strong] * np.conj(ref[strong]))
        acc[precision] = dict(rel_error=rel, rms_phase_rad=float(np.sqrt(np.mean(ph ** 2))),
                              max_phase_rad=float(np.abs(ph).max()),
                              vs="oracle-reference" if mode == 0 else "oracle-fp64",
                              pixels=int(subset.size))
        log(f"{precision} accuracy vs {acc[precision]['vs']}: rel {rel:.2e}, "
            f"rms phase {acc[precision]['rms_phase_rad']:.2e} rad")
    return acc


def host_cpu():
    """Host description for the CPU baseline: logical CPUs (nproc), the CPU
    model and the cores per socket and sockets from lscpu."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True,
                                           check=True).stdout.strip())
    except Exception:  # noqa: BLE001 - description only
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Model name", "Core(s) per socket", "Socket(s)", "Thread(s) per core"):
                info[k.lower().replace("(s)", "s").replace(" ", "_")] = v.strip()
    except Exception:  # noqa: BLE001 - description only
        pass
    return info


def cpu_threads():
    """OpenMP threads of the CPU baseline: the CPUs this process may run on
    (nproc), at most 16 (the host cores that go with one GPU of an 8-GPU,
    128-core node)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(args):
    """Time the CPU baseline in a child process (``--cpu-baseline-child``)
    that starts its OpenMP runtime with OMP_PROC_BIND=close and one thread per
    CPU.  Binding is an OpenMP runtime setting, fixed when the runtime starts
    and applied to the starting thread too: set in this process it would pin
    the main thread to one CPU, and with it every thread the library's host
    code starts from it."""
    import subprocess
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_NUM_THREADS=str(cpu_threads()))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child",
           "--cpu-seconds", str(args.cpu_seconds)]
    res = subprocess.run(cmd, env=env, capture_output=True, text=True)
    sys.stderr.write(res.stderr)
    if res.returncode != 0:
        raise RuntimeError(f"CPU baseline process failed ({res.returncode})")
    cpu = json.loads(res.stdout.strip().splitlines()[-1])
    log(f"cpu baseline: {cpu['value']:.3e} updates/s on {cpu['cores']} threads "
        f"({cpu['host']})")
    return cpu


def cpu_baseline_child(args):
    """The oracle in reference mode (the CPU restatement of the reference
    backprojector, OpenMP over pixels) on a band of output rows at the centre
    of the config-1 grid, whose pulses are simulated here from the same seed:
    one warm-up run, then the median of three timed runs, the band sized so
    that one run takes about a quarter of --cpu-seconds.  Prints one JSON
    line."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from isce3_amd import sim
    # the parent's thread count (here the OpenMP runtime has bound this thread
    # to one CPU, so this process's own affinity no longer shows the CPUs)
    nth = int(os.environ.get("OMP_NUM_THREADS", "0")) or cpu_threads()
    sc = sim.make_scene("full")
    raw_np = sim.simulate(sc)
    W = sc.out_grid.width
    row0 = sc.out_grid.length // 2

    def run(rows):
        sub = np.arange(row0 * W, (row0 + rows) * W, dtype=np.int64)
        t0 = time.perf_counter()
        _, _, diag, _ = oracle.backproject(sc.out_geometry(), raw_np, sc.in_geometry(),
                                           sc.dem, sc.fc, sc.ds, sc.kernel,
                                           sc.dry_tropo_model, subset=sub, nthreads=nth,
                                           want_diag=True)
        dt = time.perf_counter() - t0
        return dt, int(sum(d.kstop - d.kstart for d in diag if d.status == 0))

    dt1, _ = run(1)  # sizing run
    rows = int(max(1, min(sc.out_grid.length - row0, round(0.25 * args.cpu_seconds / dt1))))
    run(rows)        # warm-up at the timed size
    times = []
    upd = 0
    for _ in range(3):
        dt, upd = run(rows)
        times.append(dt)
    t_med = float(np.median(times))
    host = host_cpu()
    cpu = {"value": upd / t_med, "unit": "pixel-pulse updates/s", "cores": nth,
           "kind": "port",
           "sample": f"output rows {row0}-{row0 + rows - 1} of the {sc.out_grid.length} x {W} "
                     f"grid ({upd} updates), oracle reference mode, {nth} OpenMP threads "
                     f"(OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', '')}), median of 3 "
                     f"after a warm-up: {', '.join(f'{t:.2f}' for t in times)} s",
           "host": host}
    print(json.dumps(cpu))


# ------------------------------------------------------------ batched tiles
def device_name(torch):
    try:
        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:  # noqa: BLE001 - naming only
        return "gpu"


def run_batch(args, torch, np, dev, tuner):
    from isce3_amd import autotune, sim
    from isce3_amd.focus import backproject, backproject_blocks
    t0 = time.time()
    sc = sim.make_scene("swath")
    _, raw = make_raw(torch, sim, sc, dev)
    tiles = sim.random_tiles(sc.out_grid, 64, 64, 1024, (0.5, 2.0), seed=64)
    px = sum(t.length * t.width for t in tiles)
    log(f"batch: swath raw {tuple(raw.shape)} simulated in {time.time() - t0:.1f}s; "
        f"64 tiles, {px} pixels")
    outs = [torch.empty(t.shape, dtype=torch.complex64, device=dev) for t in tiles]
    geoms = [sc.out_geometry(t) for t in tiles]

    def batched(precision, options):
        st = {}
        backproject_blocks(outs, geoms, raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds, sc.kernel,
                           sc.dry_tropo_model, precision=precision, options=options, stats=st)
        return st

    def sequential(precision, options):
        upd = 0
        for o, g in zip(outs, geoms):
            st = {}
            backproject(o, g, raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds, sc.kernel,
                        sc.dry_tropo_model, precision=precision, options=options, stats=st)
            upd += st["updates"]
        return {"updates": upd}

    key = autotune.workload_key(device_name(torch), "split", 8, px,
                                float(np.mean([t.width for t in tiles])))
    t0 = time.time()

    def device_ms(options):
        st = batched("split", options)
        return st["geometry_ms"] + st["sum_ms"]

    best = tuner.tune(key, "split", device_ms, force=True)
    n_cand = len(tuner.trials.get(key, []))
    log(f"batch: autotuned {best} over {n_cand} candidates in {time.time() - t0:.1f}s")

    def timed(fn, precision, options):
        for _ in range(args.warmup):
            fn(precision, options)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = fn(precision, options)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        return {"ms": dt * 1e3, "updates": st["updates"], "updates_per_s": st["updates"] / dt}

    res = {"split_batched": timed(batched, "split", best),
           "plain_sequential": timed(sequential, "plain", None),
           "split_sequential": timed(sequential, "split", None),
           "plain_batched": timed(batched, "plain", None)}
    for k, v in res.items():
        log(f"batch {k}: {v['ms']:.1f} ms, {v['updates_per_s']:.3e} updates/s")
    return {"workload": "64 tiles (side log-U[64,1024] px, aspect U[0.5,2]) of one resident "
                        "8192x16384 c64 raw set",
            "pixels": int(px), "tile_px_min": int(min(t.length * t.width for t in tiles)),
            "tile_px_max": int(max(t.length * t.width for t in tiles)),
            "options": best, "tuning_candidates": n_cand, **res,
            "speedup_vs_plain_sequential": res["plain_sequential"]["ms"] / res["split_batched"]["ms"],
            "speedup_vs_split_sequential": res["split_sequential"]["ms"] / res["split_batched"]["ms"]}


# ----------------------------------------------------------------- serving
def run_serve(args, torch, np, dev, tuner):
    from isce3_amd import autotune, serve, sim
    t0 = time.time()
    sets, scenes = [], []
    for seed in (41, 42, 43):
        sc = sim.make_scene("swath", seed=seed)
        _, raw = make_raw(torch, sim, sc, dev)
        scenes.append(sc)
        sets.append(serve.RawSet(raw, sc.in_geometry(), sc.orbit, sc.doppler, sc.dem, sc.fc,
                                 sc.ds, sc.kernel, sc.dry_tropo_model))
    log(f"serve: 3 raw sets {tuple(sets[0].signal.shape)} simulated in {time.time() - t0:.1f}s")
    dname = device_name(torch)

    def key_of(batch):
        px = sum(r.pixels for r in batch)
        mw = float(np.mean([r.grid.width for r in batch]))
        return autotune.workload_key(dname, "split", 8, px, mw)

    def split_options(set_id, batch):
        return tuner.options_for(key_of(batch), "split")

    def make_requests(n, seed, set_id=None):
        rng = np.random.Generator(np.random.PCG64(seed))
        reqs = []
        for i in range(n):
            sid = int(rng.integers(0, len(sets))) if set_id is None else set_id
            (tile,) = sim.random_tiles(scenes[sid].out_grid, 1, 32, 2048, (1.0, 1.0),
                                       seed=seed * 100003 + i)
            reqs.append(serve.Request(sid, tile, tag=i))
        return reqs

    # tune the split launch options on three representative batch classes of
    # the request mix (1, 4 and 16 requests of one set)
    t0 = time.time()
    current = {}
    with serve.Server(sets, precision="split", workers=1, budget_bytes=args.serve_budget,
                      options=lambda s, b: current) as srv:
        for nb in (1, 4, 16):
            proto = make_requests(nb, seed=7000 + nb, set_id=0)

            def run(opts, proto=proto):
                current.clear()
                current.update(opts)
                n0 = len(srv.batches)
                fresh = srv.submit_many([serve.Request(0, r.grid) for r in proto])
                for r in fresh:
                    r.wait()
                return sum(b["step_ms"] for b in srv.batches[n0:])

            tuner.tune(key_of(proto), "split", run, force=True)
    log(f"serve: tuned 3 batch classes in {time.time() - t0:.1f}s")

    rows = []
    for rate in args.rates:
        n = max(20, min(args.serve_max_requests, int(args.serve_seconds * rate)))
        row = {"rate_per_s": rate, "requests": n}
        for label, precision, policy in (("split_batched", "split", "batch"),
                                         ("plain_fcfs", "plain", "fcfs")):
            reqs = make_requests(n, seed=int(rate))
            opts = split_options if precision == "split" else {}
            with serve.Server(sets, precision=precision, policy=policy, workers=3,
                              budget_bytes=args.serve_budget, options=opts) as srv:
                srv.submit(serve.Request(0, reqs[0].grid)).wait()  # warm contexts
                n0 = len(srv.batches)
                res = serve.poisson_load(srv, reqs, rate, seed=int(rate) + 1)
                batches = srv.batches[n0:]
                upd = sum(b.get("updates", 0) for b in batches)
                res.update(updates=upd, updates_per_s=upd / res["makespan_s"],
                           batches=len(batches),
                           mean_batch=float(np.mean([b["requests"] for b in batches])),
                           peak_bytes=srv.peak_in_use)
            row[label] = res
            log(f"serve {rate:g}/s {label}: p50 {res['p50_ms']:.1f} ms, p99 {res['p99_ms']:.1f} ms,"
                f" {res['updates_per_s']:.3e} updates/s, {res['batches']} batches")
        rows.append(row)
    return {"workload": "3 resident 8192x16384 c64 raw sets; Poisson tile requests, square side "
                        "log-U[32,2048] px, random set and position; "
                        f"{args.serve_budget / 1e9:.0f} GB admission budget",
            "rates": rows}


# ------------------------------------------------------- range compression
def run_rangecomp(args, torch, np, dev):
    """Range compression of 16384 pulses x 15000 samples (complex64) with a
    961-sample linear chirp, "same" mode, inputs resident in HBM: one
    FFT-convolution launch (the compressor's plan: overlap-save blocks of
    2^12 points for this shape).  HBM-bound: the roofline counts the pulses
    read and the compressed pulses written."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import rangecomp as rco
    from isce3_amd.focus import RangeComp, form_linear_chirp
    na, nr = 16384, 15000
    chirp = form_linear_chirp(0.6e12, 40e-6, 24e6)
    rc = RangeComp(chirp, nr, maxbatch=na, mode=RangeComp.Mode.Same)
    gen = torch.Generator(device=dev)
    gen.manual_seed(11)
    x = torch.randn((na, nr), dtype=torch.complex64, device=dev, generator=gen)
    y = torch.empty((na, rc.output_size), dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream()
    for _ in range(max(args.warmup, 1)):
        rc.rangecompress(y, x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        rc.rangecompress(y, x)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    nbytes = (na * nr + na * rc.output_size) * 8
    # accuracy on a sample of pulses against the FP64 oracle
    rows = [0, 1, na // 2, na - 1]
    xs = x[rows].cpu().numpy()
    ref = rco.rangecompress(chirp, xs, rco.SAME, "fp64")
    rel = float(np.linalg.norm(y[rows].cpu().numpy() - ref) / np.linalg.norm(ref))
    # CPU baseline: the oracle in complex64 (numpy's single-threaded FFTs)
    nb = 256
    xb = xs = x[:nb].cpu().numpy()
    t0 = time.perf_counter()
    rco.rangecompress(chirp, xb, rco.SAME, "fp32")
    cpu_s = time.perf_counter() - t0
    res = {"workload": f"{na} pulses x {nr} samples c64, {len(chirp)}-sample chirp, same mode",
           "fft_size": rc.fft_size, "plan": list(rc.plan), "ms": ms,
           "pulses_per_s": na / (ms * 1e-3),
           "kernel": f"rc::conv_kernel<{int(np.log2(rc.plan[1]))}>",
           "roofline": {"bound": "hbm", "achieved": nbytes / (ms * 1e-3) / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "frac": nbytes / (ms * 1e-3) / 8e12,
                        "traffic": stage_traffic("rangecomp")},
           "rel_error_vs_fp64_oracle": rel,
           "cpu_baseline": {"value": nb / cpu_s, "unit": "pulses/s", "cores": 1, "kind": "port",
                            "sample": f"{nb} pulses, numpy complex64 FFTs"}}
    log(f"rangecomp: {ms:.3f} ms for {na} pulses, {res['roofline']['achieved']:.0f} GB/s, "
        f"rel err {rel:.2e}, CPU {nb / cpu_s:.0f} pulses/s")
    rc.close()
    return res


# ------------------------------------------------------- RSLC block finishing
def run_rslc(args, torch, np, dev):
    """Finishing of a 16384 x 16384 complex64 focused block resident in HBM:
    range-correction product, block statistics and complex32 encoding in one
    pass (the per-block work of the RSLC writer).  HBM-bound: the roofline
    counts 8 B read + 8 B written back + 4 B encoded per sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import rslc as ro
    from isce3_amd.rslc import BlockFinisher
    n = 16384
    gen = torch.Generator(device=dev)
    gen.manual_seed(13)
    z = torch.randn((n, n), dtype=torch.complex64, device=dev, generator=gen)
    ph = torch.rand((n,), device=dev, generator=gen, dtype=torch.float64) * 6.283185307179586
    rc = torch.polar(torch.ones_like(ph), ph).to(torch.complex64)
    f = BlockFinisher(n * n)
    stream = torch.cuda.current_stream()
    for _ in range(max(args.warmup, 1)):
        enc, st = f.finish(z, range_cor=rc, data_type="complex32")
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        enc, st = f.finish(z, range_cor=rc, data_type="complex32")
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = sum(times) / len(times)
    nbytes = n * n * (8 + 8 + 4)
    # parity on a band of rows: one more pass from a saved copy, against the oracle
    rows = slice(n // 2, n // 2 + 64)
    zb = z[rows].cpu().numpy()
    zo, enc_o, ex = ro.finish_block(zb, rc.cpu().numpy(), None, "complex32")
    enc_b, stb = BlockFinisher(zb.size).finish(zb.copy(), range_cor=rc.cpu().numpy(),
                                               data_type="complex32")
    bit_exact = bool(np.array_equal(enc_b.view(np.uint16), enc_o.view(np.uint16)))
    # CPU baseline: the oracle (numpy, one thread) on 1024 rows
    nb = 1024
    zc = z[:nb].cpu().numpy()
    rcc = rc.cpu().numpy()
    t0 = time.perf_counter()
    ro.finish_block(zc, rcc, None, "complex32")
    cpu_s = time.perf_counter() - t0
    res = {"workload": f"{n}x{n} c64 focused block: range correction, StatsRealImag, complex32",
           "ms": ms, "samples_per_s": n * n / (ms * 1e-3),
           "kernel": "rslc::finish_kernel (+ one-workgroup merge_kernel)",
           "roofline": {"bound": "hbm", "achieved": nbytes / (ms * 1e-3) / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "frac": nbytes / (ms * 1e-3) / 8e12,
                        "traffic": stage_traffic("rslc")},
           "bit_exact_vs_oracle": bit_exact, "n_valid": st.n_valid,
           "cpu_baseline": {"value": nb * n / cpu_s, "unit": "samples/s", "cores": 1,
                            "kind": "port", "sample": f"{nb} x {n} samples, numpy oracle"}}
    log(f"rslc: {ms:.3f} ms per {n}x{n} block, {res['roofline']['achieved']:.0f} GB/s, "
        f"bit-exact {bit_exact}, CPU {nb * n / cpu_s:.3g} samples/s")
    f.close()
    del z, enc
    torch.cuda.empty_cache()
    return res


# ------------------------------------------------------- presumming
def run_presum(args, torch, np, dev):
    """BLU resampling of 16384 dithered-PRF pulses x 16384 range bins
    (complex64, resident in HBM; two sub-swaths with a 20-bin blanking gap
    walking across the swath) onto the uniform grid: host weight solves per
    (output pulse, mask pattern), then one HIP kernel for all output samples.
    HBM roofline of the kernel: 8 B read + 8 B written per output sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import presum as po
    from isce3_amd.core import LUT2d
    from isce3_amd.presum import resample
    from isce3_amd.product import RadarGridParameters
    from isce3_amd.sim import EPOCH, SPEED, straight_orbit
    nt, nr, prf, gap = 16384, 16384, 1520.0, 20
    rng = np.random.default_rng(17)
    t = np.sort((np.arange(nt) - nt / 2) / prf + rng.uniform(-0.15, 0.15, nt) / prf)
    r0, dr = 850e3, 6.25
    grid = RadarGridParameters(t[0], 0.2384, prf, r0, dr, "right", nt, nr, EPOCH)
    g = 200 + (np.arange(nt) * 997) % (nr - 400)
    sw = np.zeros((2, nt, 2), np.int64)
    sw[0, :, 1] = g
    sw[1, :, 0], sw[1, :, 1] = g + gap, nr
    xs = r0 + dr * np.array([-10.0, nr + 10.0])
    ys = t[0] + np.array([-1.0, 12.0])
    lut = LUT2d(xs, ys, 50 + 20.0 * (ys[:, None] - t[0]) + 1e-4 * (xs[None, :] - r0))
    orbit = straight_orbit()
    gen = torch.Generator(device=dev)
    gen.manual_seed(19)
    raw = torch.randn((nt, nr), dtype=torch.complex64, device=dev, generator=gen)
    stream = torch.cuda.current_stream()
    # each call's result is dropped before the next call, as a pipeline would,
    # so that every call reuses the same output block of torch's allocator
    out = None
    for _ in range(max(args.warmup, 1)):
        out = None
        out = resample(raw, t, grid, sw, orbit, lut)
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        out = None
        t0 = time.perf_counter()
        out = resample(raw, t, grid, sw, orbit, lut)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    ms = sum(times) / len(times)
    fd = lambda tt, r: 50 + 20.0 * (tt - t[0]) + 1e-4 * (r - r0)  # noqa: E731
    # parity on 16 output pulses x the first 1024 bins against the oracle (complex128)
    rows = np.arange(nt // 2, nt // 2 + 16)
    ref = po.resample(raw[:, :1024].cpu().numpy(), t, prf, r0, dr, np.minimum(sw, 1024),
                      lambda tt: SPEED, fd, rows=rows)
    got = out[nt // 2:nt // 2 + 16, :1024].cpu().numpy()
    rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    # CPU baseline: the oracle restatement (numpy, one thread) on 64 output pulses
    nb = 64
    t0 = time.perf_counter()
    po.resample(raw[:nb + 16].cpu().numpy(), t, prf, r0, dr, sw, lambda tt: SPEED, fd,
                rows=np.arange(nb))
    cpu_s = time.perf_counter() - t0
    nbytes = nt * nr * 16
    res = {"workload": f"{nt} dithered pulses x {nr} bins c64, 2 sub-swaths, {gap}-bin walking "
                       "gap, L = 12 m, Doppler plane",
           "ms_per_call": ms, "samples_per_s": nt * nr / (ms * 1e-3),
           "kernel": "presum::resample_kernel (host weights included in ms_per_call)",
           "roofline_call": {"bound": "hbm", "achieved": nbytes / (ms * 1e-3) / 1e9,
                             "peak": 8000.0, "unit": "GB/s",
                             "frac": nbytes / (ms * 1e-3) / 8e12,
                             "traffic": stage_traffic("presum")},
           "rel_error_vs_oracle": rel,
           "cpu_baseline": {"value": nb * nr / cpu_s, "unit": "output samples/s",
                            "cores": 1, "kind": "port",
                            "sample": f"{nb} output pulses x {nr} bins, numpy oracle"}}
    log(f"presum: {ms:.2f} ms per {nt}x{nr} resample, rel err {rel:.2e}, "
        f"CPU {nb * nr / cpu_s:.3g} samples/s")
    del raw, out
    torch.cuda.empty_cache()
    return res


def split_kernel_name(opts) -> str:
    """The coherent-sum kernel the split step ran with these launch options
    (kernel width 8)."""
    staged = bool((opts or {}).get("staged", 0))
    lanes = (opts or {}).get("lane_rows", 0)
    pair = not staged and lanes in (0, 2)
    return (f"sum_split_kernel<8, {'true' if staged else 'false'}, {'true' if pair else 'false'}>"
            + (" (two output lines per lane)" if pair else ""))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sample-pixels", type=int, default=2048,
                    help="pixels checked against the oracle")
    ap.add_argument("--options", type=json.loads, default=None,
                    help="fixed launch options (JSON) for the full block instead of the "
                         "autotuner's, e.g. '{\"tile_rows\": 4}'")
    ap.add_argument("--tune-cache", default=None, metavar="JSON",
                    help="file the autotuner keeps its choices in (default: "
                         "$ISCE3_AMD_TUNE_CACHE or ~/.cache/isce3_amd/autotune.json)")
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch options for the full block (profiling runs)")
    ap.add_argument("--only", choices=["plain", "split"], default=None,
                    help="time one precision of the full block only (profiling runs)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-rangecomp", action="store_true")
    ap.add_argument("--only-rangecomp", action="store_true",
                    help="time the range compressor only (profiling runs)")
    ap.add_argument("--no-rslc", action="store_true")
    ap.add_argument("--no-presum", action="store_true")
    ap.add_argument("--only-presum", action="store_true",
                    help="time the presum resampler only (profiling runs)")
    ap.add_argument("--only-rslc", action="store_true",
                    help="time the RSLC block finisher only (profiling runs)")
    ap.add_argument("--rates", type=lambda s: [float(x) for x in s.split(",")],
                    default=[20.0, 50.0, 100.0, 400.0],
                    help="request rates (the last default one saturates the GPU)")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="requests per rate = rate x this (at least 20, at most "
                         "--serve-max-requests)")
    ap.add_argument("--serve-max-requests", type=int, default=400)
    ap.add_argument("--serve-budget", type=float, default=200e9)
    ap.add_argument("--cpu-baseline-child", action="store_true",
                    help=argparse.SUPPRESS)  # the CPU baseline's own process
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of this many MiB first, as the "
                         "known-size reference for HBM byte counters")
    args = ap.parse_args()
    if args.cpu_baseline_child:
        cpu_baseline_child(args)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; runs on several GPUs are "
                          "not part of it (output tiles shard with no exchange step)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from isce3_amd import autotune

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    dev = "cuda:0"
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"calibration copy of {args.calibrate_copy} MiB done")

    if args.only_rangecomp:
        print(json.dumps({"rangecomp": run_rangecomp(args, torch, np, dev)}))
        return
    if args.only_rslc:
        print(json.dumps({"rslc": run_rslc(args, torch, np, dev)}))
        return
    if args.only_presum:
        print(json.dumps({"presum": run_presum(args, torch, np, dev)}))
        return

    tuner = autotune.Autotuner(args.tune_cache)
    sc, raw_host, results = run_full(args, torch, np, dev, tuner)
    raw_np = raw_host.numpy()
    acc = accuracy(args, np, sc, raw_np, results)
    if args.only:
        r = results[args.only]
        print(json.dumps({"only": args.only, "ms_per_step": r["ms_per_step"],
                          "sum_ms": r["sum_ms"], "geometry_ms": r["geometry_ms"],
                          "updates": r["updates"], "updates_per_s": r["updates_per_s"],
                          "accuracy": acc.get(args.only)}))
        return
    cpu = None if args.no_cpu else cpu_baseline(args)
    del raw_host, raw_np

    batch = None if args.no_batch else run_batch(args, torch, np, dev, tuner)
    torch.cuda.empty_cache()
    serve_res = None if args.no_serve else run_serve(args, torch, np, dev, tuner)
    torch.cuda.empty_cache()
    rc_res = None if args.no_rangecomp else run_rangecomp(args, torch, np, dev)
    torch.cuda.empty_cache()
    rslc_res = None if args.no_rslc else run_rslc(args, torch, np, dev)
    presum_res = None if args.no_presum else run_presum(args, torch, np, dev)

    sp, pl = results["split"], results["plain"]
    achieved = sp["updates"] * FLOPS_PER_UPDATE / (sp["sum_ms"] * 1e-3) / 1e12
    latency = None
    if serve_res:
        latency = {f"{r['rate_per_s']:g}/s": {"p50_ms": r["split_batched"]["p50_ms"],
                                              "p99_ms": r["split_batched"]["p99_ms"]}
                   for r in serve_res["rates"]}
    line = {
        "metric": "pixel-pulse updates/s (split precision) + p50/p99 tile latency; split vs plain, 1 GPU",
        "value": sp["updates_per_s"],
        "unit": "pixel-pulse updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sp["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (split)",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "16384x16384 c64 pulses resident in HBM -> 4096x4096 c64 output, "
                               "Knab 8/0.8/2048, ds=6 m, 200 targets + clutter, fractal DEM 1024^2",
                   "updates_per_step": sp["updates"], "parallelism": "1 GPU"},
        "plain": {"updates_per_s": pl["updates_per_s"], "ms_per_step": pl["ms_per_step"],
                  "sum_ms": pl["sum_ms"], "geometry_ms": pl["geometry_ms"],
                  "options": pl["options"], "dtype": "f64", "accuracy": acc.get("plain")},
        "split": {"updates_per_s": sp["updates_per_s"], "ms_per_step": sp["ms_per_step"],
                  "sum_ms": sp["sum_ms"], "geometry_ms": sp["geometry_ms"],
                  "options": sp["options"], "accuracy": acc.get("split")},
        "split_over_plain": sp["updates_per_s"] / pl["updates_per_s"],
        "roofline": {"bound": "valu", "kernel": split_kernel_name(sp["options"]),
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": hbm_bytes_per_launch()},
        "latency": latency,
        "batch": batch,
        "serve": serve_res,
        "rangecomp": rc_res,
        "rslc": rslc_res,
        "presum": presum_res,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["split_over_cpu"] = sp["updates_per_s"] / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
