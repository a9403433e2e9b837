"""Launch-option autotuner of the coherent-sum kernels.

The tunable options are those of ``bp_options`` (include/isce3_amd_backproject.h):

    tile_rows      output lines per 64-pixel wavefront tile (1, 2, 4, 8)
    order          tile walk of the persistent split kernel (1 row-major,
                   2 column-major)
    chunk          pulses per compensated FP32 partial sum and per phase-step
                   interval (split; at most 64)
    blocks_per_cu  resident 512-thread workgroups per CU (0 = occupancy limit)
    staged         split taps gathered from global memory (0) or read from
                   pulse windows staged in LDS (1)
    lane_rows      output lines per lane of the split sum (2: adjacent lines
                   share one gather; staged 0 only -- with staged 1 the call
                   raises InvalidArgument and the candidate scores as
                   infinitely slow)

The plain kernel only has ``tile_rows``.  A tuning run times candidates with
the device-side HIP-event times that every call reports (geometry + sum),
by coordinate descent from the defaults, and keeps the best per workload key
(device, precision, kernel width, a pixel-count bucket and a tile-width
bucket) in a JSON cache so later processes reuse it without re-timing.
"""
from __future__ import annotations

import json
import math
import os
import threading
from typing import Callable, Dict, Iterable, Optional

DEFAULTS = {"tile_rows": 8, "order": 2, "chunk": 32, "blocks_per_cu": 0, "staged": 0,
            "lane_rows": 2}
SPACE = {
    "split": {"tile_rows": (1, 2, 4, 8), "order": (1, 2), "chunk": (16, 32, 64),
              "blocks_per_cu": (0, 1, 2), "staged": (0, 1), "lane_rows": (1, 2)},
    "plain": {"tile_rows": (1, 2, 4, 8)},
}


def default_cache_path() -> str:
    env = os.environ.get("ISCE3_AMD_TUNE_CACHE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "isce3_amd", "autotune.json")


def workload_key(device_name: str, precision: str, kernel_width: int, pixels: int,
                 mean_width: float) -> str:
    """Key of a workload class: pixel count and mean tile width in powers of 2."""
    pb = int(math.floor(math.log2(max(pixels, 1))))
    wb = int(math.floor(math.log2(max(mean_width, 1.0))))
    return f"{device_name}|{precision}|w{kernel_width}|px2^{pb}|tw2^{wb}"


def _split_key(key: str):
    head, px, tw = key.rsplit("|", 2)
    if not (px.startswith("px2^") and tw.startswith("tw2^")):
        raise ValueError(key)
    return head, int(px[4:]), int(tw[4:])


class Autotuner:
    """Times option candidates through a caller-supplied run function and
    caches the fastest per workload key."""

    def __init__(self, cache_path: Optional[str] = None, repeats: int = 2):
        self.cache_path = cache_path or default_cache_path()
        self.repeats = max(1, int(repeats))
        self._lock = threading.Lock()
        self._cache: Dict[str, dict] = {}
        self.trials: Dict[str, list] = {}
        self._load()

    def _load(self):
        try:
            with open(self.cache_path) as f:
                data = json.load(f)
            if isinstance(data, dict):
                self._cache = {k: v for k, v in data.items() if isinstance(v, dict)}
        except (OSError, ValueError):
            self._cache = {}

    def _save(self):
        try:
            os.makedirs(os.path.dirname(self.cache_path) or ".", exist_ok=True)
            tmp = self.cache_path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self._cache, f, indent=1, sort_keys=True)
            os.replace(tmp, self.cache_path)
        except OSError:
            pass  # a read-only home only loses the cache, not the result

    def lookup(self, key: str) -> Optional[dict]:
        with self._lock:
            hit = self._cache.get(key)
            return dict(hit["options"]) if hit else None

    def nearest(self, key: str) -> Optional[dict]:
        """Options of the cached key closest to `key` (same device, precision
        and kernel width; nearest pixel and tile-width buckets), or None."""
        try:
            head, px, tw = _split_key(key)
        except ValueError:
            return None
        best, best_d = None, None
        with self._lock:
            for k, v in self._cache.items():
                try:
                    h2, px2, tw2 = _split_key(k)
                except ValueError:
                    continue
                if h2 != head:
                    continue
                d = abs(px2 - px) + abs(tw2 - tw)
                if best_d is None or d < best_d:
                    best, best_d = v["options"], d
        return dict(best) if best is not None else None

    def options_for(self, key: str, precision: str) -> dict:
        """Cached options of `key`, else the defaults valid for `precision`."""
        hit = self.lookup(key)
        if hit is None:
            hit = self.nearest(key)
        if hit is not None:
            return hit
        return {k: v for k, v in DEFAULTS.items() if k in SPACE[precision]}

    def tune(self, key: str, precision: str, run: Callable[[dict], float],
             space: Optional[Dict[str, Iterable]] = None, force: bool = False) -> dict:
        """Coordinate descent over `space` (default SPACE[precision]).

        `run(options)` performs one call with those options and returns its
        device time in ms.  Each candidate is timed `repeats` times and scored
        by its minimum.  Returns the best options (also cached under `key`).
        """
        if not force:
            hit = self.lookup(key)
            if hit is not None:
                return hit
        space = {k: tuple(v) for k, v in (space or SPACE[precision]).items()}
        best = {k: DEFAULTS[k] for k in space}
        seen: Dict[tuple, float] = {}
        log = []

        def score(opts):
            sig = tuple(sorted(opts.items()))
            if sig not in seen:
                try:
                    seen[sig] = min(float(run(dict(opts))) for _ in range(self.repeats))
                except ValueError:  # InvalidArgument: a combination the kernels do not offer
                    seen[sig] = math.inf
                log.append((dict(opts), seen[sig]))
            return seen[sig]

        best_ms = score(best)
        for _ in range(2):
            improved = False
            for name, values in space.items():
                for v in values:
                    if v == best[name]:
                        continue
                    cand = dict(best)
                    cand[name] = v
                    ms = score(cand)
                    if ms < best_ms:
                        best, best_ms, improved = cand, ms, True
            if not improved:
                break
        with self._lock:
            self._cache[key] = {"options": best, "ms": best_ms, "candidates": len(seen)}
            self.trials[key] = log
            self._save()
        return dict(best)
