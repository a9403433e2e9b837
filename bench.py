#!/usr/bin/env python3
"""Benchmark: pixel-pulse updates/s of the backprojector on one MI355X.

Workload (one "step" = one full backprojection of one block, inputs resident
in HBM): 16384 pulses x 16384 range bins complex64 (2 GiB) of a seeded
synthetic L-band stripmap scene (200 point targets + clutter, 1024^2 fractal
DEM), focused to a 4096 x 4096 complex64 grid with a tabulated Knab kernel
(width 8, bandwidth 0.8, 2048 entries) at 6 m azimuth resolution.

`value` is the split-precision path (FP64 per-pixel reference + FP32 deltas);
the plain all-FP64-inner-loop kernel is timed on the same inputs beside it.
Accuracy of both is stated on a seeded pixel sample against the CPU oracle
(reference mode for plain, all-FP64 mode for split).  The CPU baseline is the
oracle in reference mode on a bounded sample of output rows, timed on the
host's cores.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_UPDATE = 101  # algorithmic flops per pixel-pulse update (FMA = 2)
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sample-pixels", type=int, default=2048,
                    help="pixels checked against the oracle")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--only", choices=["plain", "split"], default=None,
                    help="time one precision only (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of this many MiB first, as the "
                         "known-size reference for HBM byte counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not "
                          "part of it (output tiles shard with no exchange step)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from isce3_amd import sim
    from isce3_amd.focus import backproject

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    dev = "cuda:0"
    t0 = time.time()
    sc = sim.make_scene(args.config)
    log(f"scene {args.config}: {len(sc.targets)} scatterers ({time.time() - t0:.1f}s)")
    t0 = time.time()
    raw_host = torch.empty((sc.in_grid.length, sc.in_grid.width), dtype=torch.complex64,
                           pin_memory=True)
    sim.simulate(sc, out=raw_host.numpy(), nthreads=min(16, os.cpu_count() or 1))
    log(f"simulated {tuple(raw_host.shape)} in {time.time() - t0:.1f}s")
    raw = raw_host.to(dev)
    torch.cuda.synchronize()
    out = torch.empty(sc.out_grid.shape, dtype=torch.complex64, device=dev)
    opts = {"tile_rows": args.tile_rows} if args.tile_rows else None
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"calibration copy of {args.calibrate_copy} MiB done")

    def step(precision):
        st = {}
        backproject(out, sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                    sc.kernel, sc.dry_tropo_model, precision=precision, stats=st, options=opts)
        return st

    results = {}
    precisions = (args.only,) if args.only else ("plain", "split")
    for precision in precisions:
        for _ in range(args.warmup):
            step(precision)
        torch.cuda.synchronize()
        stats = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            stats.append(step(precision))
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        upd = stats[-1]["updates"]
        sum_ms = float(np.mean([s["sum_ms"] for s in stats]))
        geo_ms = float(np.mean([s["geometry_ms"] for s in stats]))
        results[precision] = dict(ms_per_step=dt * 1e3, updates=upd, sum_ms=sum_ms,
                                  geometry_ms=geo_ms, updates_per_s=upd / dt,
                                  out=out.cpu().numpy())
        log(f"{precision}: {dt * 1e3:.1f} ms/step, sum {sum_ms:.1f} ms, geometry {geo_ms:.1f} ms,"
            f" {upd / dt:.3e} updates/s")

    # accuracy on a seeded pixel sample against the oracle
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    raw_np = raw_host.numpy()
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
                              vs="oracle-reference" if mode == 0 else "oracle-fp64")
        log(f"{precision} accuracy vs {acc[precision]['vs']}: rel {rel:.2e}, "
            f"rms phase {acc[precision]['rms_phase_rad']:.2e} rad")

    cpu = None
    if not args.no_cpu:
        W = sc.out_grid.width
        rows = 1
        t_total = 0.0
        upd_total = 0
        row0 = sc.out_grid.length // 2
        while True:
            sub = np.arange(row0 * W, (row0 + rows) * W, dtype=np.int64)
            t0 = time.perf_counter()
            _, _, diag, _ = oracle.backproject(sc.out_geometry(), raw_np, sc.in_geometry(),
                                               sc.dem, sc.fc, sc.ds, sc.kernel,
                                               sc.dry_tropo_model, subset=sub, nthreads=nth,
                                               want_diag=True)
            dt = time.perf_counter() - t0
            t_total += dt
            upd_total += int(sum(d.kstop - d.kstart for d in diag if d.status == 0))
            row0 += rows
            if t_total >= args.cpu_seconds or row0 >= sc.out_grid.length:
                break
            rows = max(1, min(64, int(rows * max(1.0, (args.cpu_seconds - t_total) / max(dt, 1e-3)))))
        cpu = {"value": upd_total / t_total, "unit": "pixel-pulse updates/s", "cores": nth,
               "kind": "port",
               "sample": f"{upd_total} updates ({upd_total / sc.out_grid.width / 4046:.0f}-ish "
                         f"rows of 4096 px) of the same workload, oracle reference mode, "
                         f"{t_total:.1f} s"}
        log(f"cpu baseline: {cpu['value']:.3e} updates/s on {nth} threads")

    if args.only:
        r = results[args.only]
        print(json.dumps({"only": args.only, "ms_per_step": r["ms_per_step"],
                          "sum_ms": r["sum_ms"], "updates": r["updates"],
                          "updates_per_s": r["updates_per_s"]}))
        return
    sp, pl = results["split"], results["plain"]
    achieved = sp["updates"] * FLOPS_PER_UPDATE / (sp["sum_ms"] * 1e-3) / 1e12
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
        "config": {"workload": "c1/c2: 16384x16384 c64 pulses -> 4096x4096 c64, Knab 8/0.8/2048,"
                               " ds=6 m, 200 targets + clutter, fractal DEM 1024^2",
                   "updates_per_step": sp["updates"], "parallelism": "1 GPU"},
        "plain": {"updates_per_s": pl["updates_per_s"], "ms_per_step": pl["ms_per_step"],
                  "sum_ms": pl["sum_ms"], "geometry_ms": pl["geometry_ms"],
                  "dtype": "f64", "accuracy": acc["plain"]},
        "split": {"updates_per_s": sp["updates_per_s"], "ms_per_step": sp["ms_per_step"],
                  "sum_ms": sp["sum_ms"], "geometry_ms": sp["geometry_ms"],
                  "accuracy": acc["split"]},
        "split_over_plain": sp["updates_per_s"] / pl["updates_per_s"],
        "roofline": {"bound": "valu", "kernel": "sum_split_kernel<8>",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None},
        "latency": None,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["split_over_cpu"] = sp["updates_per_s"] / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
