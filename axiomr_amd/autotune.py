"""Autotuner for the render path's tunables (axr_config): raster mode
(binned / visibility buffer), bin edge (8/16/32 px), threads per bin
workgroup (64/128/256), per-wave bin-pair capacity (64..1024, binned mode)
and the one-lane footprint threshold.  None of them changes a pixel; they
only move time between the kernels.

Coordinate descent: raster mode x bin edge x workgroup size first (18
points), then the bin capacity (binned mode only), then the footprint
threshold, each point timed as the median of a few synchronised runs of the
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
DEFAULT_CACHE = Path(__file__).resolve().parent / "tuning" / "mi355x.json"


def time_config(ctx, frame: Callable[[], None], reps: int = 3, inner: int = 2) -> float:
    """Median seconds of `inner` frames, over `reps` runs, after a warm-up."""
    frame()
    ctx.sync()
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
    start = dict(best)
    for mode in MODES:
        for tile in TILES:
            for nt in THREADS:
                cfg = {**start, "raster_mode": mode, "tile_size": tile, "raster_threads": nt}
                if cfg == start:
                    continue
                t = run(cfg)
                if t < best_t:
                    best, best_t = cfg, t
    if best.get("raster_mode", 0) == 0:
        for cap in CAPACITIES:
            cfg = {**best, "bin_capacity": cap}
            if cfg != best:
                t = run(cfg)
                if t < best_t:
                    best, best_t = cfg, t
    for fp in FOOTPRINTS:
        cfg = {**best, "small_footprint": fp}
        if cfg != best:
            t = run(cfg)
            if t < best_t:
                best, best_t = cfg, t
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
