"""Autotuner for the render path's tunables (axr_config): raster mode
(binned / visibility buffer), bin edge (8/16/32 px), threads per bin
workgroup (64/128/256), per-wave bin-pair capacity (64..1024, binned mode)
and the one-lane footprint threshold.  None of them changes a pixel; they
only move time between the kernels.

Coordinate descent, two rounds: raster mode x bin edge x workgroup size
(18 points), then the bin capacity (binned mode only; the smallest within
1 % of the fastest, since it sizes memory), then the footprint threshold,
each point timed as the median of a few synchronised runs of the
caller's frame function.  Results are cached in a JSON file keyed by device
and workload name.
"""
from __future__ import annotations

import json
import time
from pathlib import Path
from typing import Callable, Dict, Optional

MODES = (0, 1)          # AXR_RASTER_BINNED, AXR_RASTER_VISIBILITY
TILES = (8, 16, 32)
THREADS = (64, 128, 256)
CAPACITIES = (64, 128, 256, 512, 1024)
FOOTPRINTS = (8, 16, 32, 64, 128, 256)
CAPACITY_SLACK = 0.01   # a larger bin capacity must be this much faster to be kept
DEFAULT_CACHE = Path(__file__).resolve().parent / "tuning" / "mi355x.json"


def _warm(ctx, frame: Callable[[], None], n: int, tries: int = 8) -> None:
    """n frames and a wait, again while a frame arena reports an overflow (it
    is grown at the wait; the frames since the previous wait were not drawn)."""
    from ._native import AXR_ERR_OVERFLOW, AxrError
    for attempt in range(tries):
        for _ in range(n):
            frame()
        try:
            ctx.sync()
            return
        except AxrError as e:
            if e.code != AXR_ERR_OVERFLOW or attempt == tries - 1:
                raise


def time_config(ctx, frame: Callable[[], None], reps: int = 3, inner: int = 2) -> float:
    """Median seconds of `inner` frames, over `reps` runs, after a warm-up of
    as many frames (asynchronous frames overlap, so a short run would time
    the pipeline filling instead of its steady state)."""
    _warm(ctx, frame, inner)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for _ in range(inner):
            frame()
        ctx.sync()
        ts.append((time.perf_counter() - t0) / inner)
    ts.sort()
    return ts[len(ts) // 2]


def tune(ctx, make_frame: Callable[[], Callable[[], None]], log: Optional[Callable[[str], None]] = None,
         reps: int = 3, inner: int = 2) -> Dict:
    """`make_frame()` is called after every config change and returns a
    callable that enqueues one frame (it may rebuild batch handles, which
    bind the config at creation)."""
    best = dict(ctx.config())
    trials = []

    def run(cfg):
        ctx.set_config(**cfg)
        t = time_config(ctx, make_frame(), reps, inner)
        trials.append({**cfg, "seconds": t})
        if log:
            log(f"[autotune] {cfg} -> {t * 1e3:.3f} ms")
        return t

    best_t = run(best)
    seen = {tuple(sorted(best.items()))}

    def consider(cfg):
        nonlocal best, best_t
        key = tuple(sorted(cfg.items()))
        if key in seen:
            return
        seen.add(key)
        t = run(cfg)
        if t < best_t:
            best, best_t = cfg, t

    # two rounds: the second sweeps the launch shape again under the
    # footprint and capacity the first one chose
    for _ in range(2):
        start = dict(best)
        for mode in MODES:
            for tile in TILES:
                for nt in THREADS:
                    consider({**start, "raster_mode": mode, "tile_size": tile, "raster_threads": nt})
        if best.get("raster_mode", 0) == 0:
            for cap in CAPACITIES:
                consider({**best, "bin_capacity": cap})
            # the capacity sizes the per-wave pair regions (setup waves x
            # capacity x 16 B): take the smallest within CAPACITY_SLACK of
            # the fastest, whose edge is below the timing noise
            times = {t["bin_capacity"]: t["seconds"] for t in trials
                     if all(t[k] == v for k, v in best.items() if k != "bin_capacity")}
            for cap in sorted(times):
                if times[cap] <= best_t * (1.0 + CAPACITY_SLACK):
                    best, best_t = {**best, "bin_capacity": cap}, times[cap]
                    break
        for fp in FOOTPRINTS:
            consider({**best, "small_footprint": fp})
    ctx.set_config(**best)
    return {"best": best, "seconds": best_t, "trials": trials}


def load(workload: str, device: str = "gfx950", path: Path = DEFAULT_CACHE) -> Optional[Dict]:
    if not path.exists():
        return None
    data = json.loads(path.read_text())
    return data.get(device, {}).get(workload)


def store(workload: str, result: Dict, device: str = "gfx950", path: Path = DEFAULT_CACHE) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    data = json.loads(path.read_text()) if path.exists() else {}
    data.setdefault(device, {})[workload] = {"config": result["best"], "seconds": result["seconds"]}
    path.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
