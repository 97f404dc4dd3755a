"""Benchmark of the MI355X tiled render path.

Default workload (N=1): the plain per-frame GPU path at 1920x1080 on a
seeded, noise-displaced icosphere of 1,003,520 triangles (frequency 224),
normal-mapped Phong with a 2048^2 RGBA8 noise albedo, a 2048^2 normal map
(both bilinear) and 3 point lights, FP32 depth, BGRA8 output.  One step =
one frame = framebuffer clear + drawMesh, inputs resident in HBM.

Prints one JSON line: frames/s (value), Gtri/s (submitted triangles), the
roofline of the dominant kernel measured with HIP events on the context's
stream, and the CPU checker (the restated reference tiled pipeline, 16
threads) timed on a bounded sample of the same frame on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

METRIC = "frames/s + Gtri/s @1080p 1M-tri vs plain GPU path; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch measured with rocprofv3 --pmc (FETCH_SIZE, then
# WRITE_SIZE) on this workload, scaled by a plain-copy calibration; written
# by `python -m axiomr_amd.rocprof_summary traffic`.
TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_icosphere1080p.json"
BATCH_TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_batch256.json"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(nverts, nfaces, W, H, covered, ntex):
    """B = 12 V + 12 T + 8 W H + 4 N_tex P_cov: positions, indices, one
    colour+depth write per pixel, one texel per texture per covered pixel."""
    return 12 * nverts + 12 * nfaces + 8 * W * H + 4 * ntex * covered


def raster_bytes(covered, ntex):
    """Algorithmic bytes of one k_raster launch: per covered pixel, the depth
    read of the framebuffer test, the depth + colour write and one texel per
    texture."""
    return covered * (4 + 8 + 4 * ntex)


def resolve_bytes(W, H, covered, ntex):
    """Algorithmic bytes of one k_resolve launch, which also writes the
    frame's clear: colour + depth of every pixel and one texel per texture
    per covered pixel (the winners are tested against the clear depth, so
    no framebuffer read)."""
    return 8 * W * H + 4 * ntex * covered


def setup_bytes(nverts, nfaces):
    """Algorithmic bytes of one k_setup launch: indices + positions."""
    return 12 * nfaces + 12 * nverts


RASTER_MODE = None   # --raster-mode: overrides the tuned or default mode


def apply_config(ctx, workload, tile, use_tuned):
    """The tuned tunables for `workload` when cached (and asked for), else
    the defaults with --tile; then --raster-mode if given."""
    from axiomr_amd import autotune as AT
    cached = AT.load(workload) if use_tuned else None
    if cached:
        ctx.set_config(**cached["config"])
    elif tile:
        ctx.set_config(tile_size=tile)
    if RASTER_MODE is not None:
        ctx.set_config(raster_mode=RASTER_MODE)
    return ctx.config()


def warm(ctx, enqueue, tries=8):
    """Run `enqueue` (asynchronous frames) and wait for it.  A frame arena
    that overflowed is reported at the wait, grown and reset, and the frames
    since the previous wait were not drawn: run again until none overflows."""
    from axiomr_amd._native import AXR_ERR_OVERFLOW, AxrError
    for attempt in range(tries):
        enqueue()
        try:
            ctx.sync()
            return
        except AxrError as e:
            if e.code != AXR_ERR_OVERFLOW or attempt == tries - 1:
                raise


def stage_names(cfg, two_pass=False):
    """Kernel names of the five timed stages in the context's raster mode;
    two_pass: visibility-mode draws that resolve in two launches (an
    asynchronous draw with a clear)."""
    if cfg.get("raster_mode", 0) == 1:
        return ("k_vis", "k_clip", "k_big", "k_shade", "k_finish", "k_vary") if two_pass else \
            ("k_vis", "k_clip", "k_big", "none", "k_resolve", "k_vary")
    return ("k_setup", "k_clip", "k_scan", "k_scatter", "k_raster", "k_vary")


def stage_times(ctx, cfg, two_pass=False):
    """Average kernel time (ms) of each stage over the frames timed since
    profiling was switched on: HIP events attached to the kernels' own
    dispatches (hipExtLaunchKernel), so they read what a kernel trace reads."""
    st = ctx.stats()
    n = max(1, st.frames)
    return {name: st.stage_ms_total[i] / n for i, name in enumerate(stage_names(cfg, two_pass))
            if name != "none" and st.stage_ms_total[i] > 0}


def compare_frames(gpu, ref, gpu_ids=None):
    """Parity of a GPU frame (bgra, depth) against the checker's frame
    (bgra, depth, ids): coverage and depth bits must match exactly, colour
    to within 1 LSB per channel.  gpu_ids: the winner ids of a frame drawn
    by the same kernels into a framebuffer with an id plane."""
    gc, gd = gpu[0], gpu[1]
    rc, rd, ri = ref
    inf = np.float32(np.inf).view(np.uint32)
    cov_g, cov_r = gd.view(np.uint32) != inf, rd.view(np.uint32) != inf
    diff = np.abs(gc.astype(np.int16) - rc.astype(np.int16))
    rep = {"pixels": int(gd.size), "covered": int(cov_r.sum()),
           "coverage_mismatch": int((cov_g != cov_r).sum()),
           "depth_mismatch": int((gd.view(np.uint32) != rd.view(np.uint32)).sum()),
           "winner_mismatch": int((gpu_ids != ri).sum()) if gpu_ids is not None else None,
           "color_max": int(diff.max()) if diff.size else 0,
           "color_mean": float(diff.mean()) if diff.size else 0.0}
    return rep


def parity_ok(rep):
    return (rep["coverage_mismatch"] == 0 and rep["depth_mismatch"] == 0 and not rep.get("winner_mismatch")
            and rep["color_max"] <= 1)


def merge_parity(acc, rep):
    """Sum of per-view parity reports (colour max is the max, mean is
    pixel-weighted)."""
    if acc is None:
        return dict(rep)
    out = dict(acc)
    for k in ("pixels", "covered", "coverage_mismatch", "depth_mismatch"):
        out[k] = acc[k] + rep[k]
    if acc.get("winner_mismatch") is not None and rep.get("winner_mismatch") is not None:
        out["winner_mismatch"] = acc["winner_mismatch"] + rep["winner_mismatch"]
    out["color_max"] = max(acc["color_max"], rep["color_max"])
    out["color_mean"] = (acc["color_mean"] * acc["pixels"] + rep["color_mean"] * rep["pixels"]) / out["pixels"]
    return out


def run_frames(scene, steps, warmup, tile, use_tuned, workload="plain"):
    """The plain per-frame GPU path on one scene: `steps` frames of
    clear + asynchronous drawMesh into one framebuffer, timed between two
    stream synchronisations; the last timed frame is read back.  Then as
    many frames again with events on each kernel's dispatch give the
    kernels' durations, and one more frame through the same kernels into a
    framebuffer with an id plane gives the winners."""
    import axiomr_amd as axr
    from axiomr_amd import render as R
    from axiomr_amd import resident as H

    ctx = axr.Context(0)
    cfg = apply_config(ctx, workload, tile, use_tuned)
    gs = H.GpuScene(ctx, scene)
    fb = R.Framebuffer(ctx, scene.width, scene.height)
    params = H.params_of(scene, flags=1)   # AXR_DRAW_ASYNC
    L = ctx._lib
    bgra = H.BLACK.packed_bgra()
    inf = float("inf")
    pp = C.byref(params)

    def frame(target=fb):
        L.axr_fb_clear(target.handle, bgra, inf)
        L.axr_draw(ctx.handle, target.handle, gs.mesh.handle, pp)

    # first frame synchronously: sizes the frame arena, gives coverage
    fb.clear(H.BLACK)
    p0 = H.params_of(scene)
    ctx.check(L.axr_draw(ctx.handle, fb.handle, gs.mesh.handle, C.byref(p0)))
    _, depth, _ = fb.read()
    covered = int(np.isfinite(depth).sum())

    warm(ctx, lambda: [frame() for _ in range(max(1, warmup))])
    t0 = time.perf_counter()
    for _ in range(steps):
        frame()
    t_enq = time.perf_counter() - t0
    ctx.sync()
    dt = time.perf_counter() - t0
    last = fb.read()
    # kernel times: the same frames again with events on each kernel's
    # dispatch (a dispatch that carries events runs slower end to end, so
    # these frames are not the ones `value` times)
    ctx.set_profiling(True)
    for _ in range(steps):
        frame()
    ctx.sync()
    stages = stage_times(ctx, cfg, two_pass=True)
    ctx.set_profiling(False)
    fbi = R.Framebuffer(ctx, scene.width, scene.height, with_ids=True)
    frame(fbi)
    ctx.sync()
    checked = fbi.read()
    fbi.close()
    in_use, arena = ctx.memory()
    fb.close()
    gs.close()
    ctx.close()
    return dict(seconds=dt, enqueue_seconds=t_enq, covered=covered, stages_ms=stages, mem_in_use=in_use,
                arena_bytes=arena, config=cfg, last=last, checked=checked)


def frame_kernel_bytes(nverts, nfaces, W, H, covered, ntex):
    """Algorithmic bytes of each kernel of the plain visibility-mode frame
    (they sum to the frame's B = 12 V + 12 T + 8 W H + 4 N_tex P_cov): k_vis
    reads positions and indices, k_shade one texel per texture per covered
    pixel, k_finish writes colour and depth of every pixel (the absorbed
    clear included); the clipper and large-triangle passes have ~none."""
    return {"k_vis": setup_bytes(nverts, nfaces), "k_clip": 0, "k_big": 0, "k_shade": 4 * ntex * covered,
            "k_finish": 8 * W * H, "k_resolve": resolve_bytes(W, H, covered, ntex)}


def run_batch(wl, reps, tile, use_tuned):
    """The batched form (256 views, 32 meshes) against the plain per-frame
    path on the same views: frames (views) per second of each.  Returns the
    measurements plus the live context and batch (`gb`), whose framebuffers
    hold the batch's last frames for the parity check; the caller closes
    both."""
    import axiomr_amd as axr
    from axiomr_amd import resident as H

    ctx = axr.Context(0)
    # plain per-frame path: clear + drawMesh for every view, queued; timed
    # under both tuned configurations, the faster one is the baseline
    plain_s, plain_cfg = None, None
    gb = None
    for which in ("plain", "batch"):
        cfg = apply_config(ctx, which, tile, use_tuned)
        if gb is None:
            gb = H.GpuBatch(ctx, wl, with_ids=False)
        n = len(wl.views)
        warm(ctx, lambda: [gb.draw_solo(i, flags=1) for i in range(n)])
        t0 = time.perf_counter()
        for _ in range(reps):
            for i in range(n):
                gb.draw_solo(i, flags=1)
        ctx.sync()
        dt = (time.perf_counter() - t0) / reps
        if plain_s is None or dt < plain_s:
            plain_s, plain_cfg = dt, cfg
    # the same views through the synchronous drop-in drawMesh (each draw
    # returns when its frame is written, as the reference's does), plain config
    apply_config(ctx, "plain", tile, use_tuned)
    for i in range(n):
        gb.draw_solo(i)
    t0 = time.perf_counter()
    for i in range(n):
        gb.draw_solo(i)
    sync_s = time.perf_counter() - t0
    cfg = apply_config(ctx, "batch", tile, use_tuned)
    gb.rebuild_batch()
    # batched: one graph launch draws (and clears) every view
    gb.batch.launch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        gb.batch.launch()
    ctx.sync()
    batch_s = (time.perf_counter() - t0) / reps
    # kernel times: the same launch sequence outside the graph, with events
    # on each kernel's dispatch
    ctx.set_profiling(True)
    for _ in range(3):
        gb.batch.launch()
    ctx.sync()
    stages = stage_times(ctx, cfg)
    ctx.set_profiling(False)
    gb.batch.launch()   # the frames the parity check reads: drawn by the graph
    ctx.sync()
    covered = 0
    for fbv in gb.fbs:
        _, d, _ = fbv.read()
        covered += int(np.isfinite(d).sum())
    return dict(views=n, plain_views_per_s=n / plain_s, batch_views_per_s=n / batch_s, speedup=plain_s / batch_s,
                sync_views_per_s=n / sync_s, speedup_vs_sync=sync_s / batch_s,
                batch_ms=batch_s * 1e3, plain_ms=plain_s * 1e3, stages_ms=stages, covered=covered,
                faces=wl.total_faces, pixels=wl.total_pixels, config=cfg, plain_config=plain_cfg, gb=gb, ctx=ctx)


def batch_kernel_bytes(wl, covered, ntex=2):
    """Algorithmic bytes per launch of the batch's kernels: k_setup reads
    every view's positions and indices (12 V + 12 T), k_raster writes colour
    and depth of the covered pixels and reads one texel per texture per
    covered pixel (the clear pass writes the rest of every view)."""
    verts = sum(wl.meshes[v.mesh][0].shape[0] for v in wl.views)
    return {"k_setup": 12 * verts + 12 * wl.total_faces, "k_raster": (8 + 4 * ntex) * covered,
            "k_clear_views": 8 * wl.total_pixels, "k_vary": 0}


SERVE_RATES = (100, 250, 500, 1000, 2500, 5000, 10000, 20000, 40000)


def run_serve(wl, rates, seconds_per_rate, lanes, budget, tile, use_tuned, burst=2000):
    """The served form: an open-loop Poisson stream at each rate, replayed
    natively (submission at the scheduled instants, results released on
    completion); p50/p99 latency = completion - scheduled arrival.  Then a
    burst (every request at t=0) for the capacity in frames/s, and the same
    burst through the plain per-frame path (clear + drawMesh per request,
    queued on one stream) as the baseline."""
    import axiomr_amd as axr
    from axiomr_amd import resident as H
    from axiomr_amd import scenes as S
    from axiomr_amd.serve import Server

    ctx = axr.Context(0)
    cfg = apply_config(ctx, "serve", tile, use_tuned) if use_tuned and _has_tuned("serve") else \
        apply_config(ctx, "plain", tile, use_tuned)
    gs = H.GpuServeScene(ctx, wl)
    faces = [f.shape[0] for _, f in wl.meshes]

    def build(views):
        return [gs.request(v) for v in views]

    srv = Server(ctx, lanes=lanes, max_batch_views=64, hbm_budget=budget)
    srv.replay(build(wl.requests(300, seed=1)), np.zeros(300))      # warm the pool and the lanes
    out = []
    for rate in rates:
        n = int(min(4000, max(200, rate * seconds_per_rate)))
        views = wl.requests(n, seed=rate)
        reqs = build(views)
        arr = S.poisson_arrivals(n, rate, seed=rate)
        st0 = srv.stats()
        lat = srv.replay(reqs, arr)
        st1 = srv.stats()
        span = float(np.max(arr + lat))
        nb = st1["batches"] - st0["batches"]
        rec = {"rate": rate, "requests": n, "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
               "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
               "max_ms": round(float(lat.max()) * 1e3, 3),
               "frames_per_s": round(n / span, 1),
               "gtri_per_s": round(sum(faces[v.mesh] for v in views) / span / 1e9, 4),
               "mean_views_per_batch": round((st1["batched_views"] - st0["batched_views"]) / max(1, nb), 2),
               "overlapped_launches": st1["launches_overlapped"] - st0["launches_overlapped"]}
        log(f"[bench] serve {rate} req/s: {rec}")
        out.append(rec)
    # capacity: a burst of `burst` requests at t=0
    views = wl.requests(burst, seed=7)
    reqs = build(views)
    t0 = time.perf_counter()
    lat = srv.replay(reqs, np.zeros(burst))
    cap_s = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    # plain per-frame path on the same burst
    import ctypes as C
    from axiomr_amd import render as R
    fbs = {}
    for v in views:
        if (v.width, v.height) not in fbs:
            fbs[(v.width, v.height)] = R.Framebuffer(ctx, v.width, v.height)
    L = ctx._lib
    bgra = H.BLACK.packed_bgra()
    plist = [gs.params(v) for v in views]
    for p in plist:
        p.flags = 1   # AXR_DRAW_ASYNC

    def plain_all():
        for v, p in zip(views, plist):
            fb = fbs[(v.width, v.height)]
            L.axr_fb_clear(fb.handle, bgra, float("inf"))
            L.axr_draw(ctx.handle, fb.handle, gs.meshes[v.mesh].handle, C.byref(p))
        ctx.sync()

    plain_all()
    t0 = time.perf_counter()
    plain_all()
    plain_s = time.perf_counter() - t0
    for fb in fbs.values():
        fb.close()
    gs.close()
    ctx.close()
    burst_faces = sum(faces[v.mesh] for v in views)
    return dict(rates=out, capacity_frames_per_s=burst / cap_s, capacity_gtri_per_s=burst_faces / cap_s / 1e9,
                burst_p50_ms=float(np.percentile(lat, 50)) * 1e3, plain_frames_per_s=burst / plain_s,
                speedup=plain_s / cap_s, stats=st, config=cfg)


def _has_tuned(workload):
    from axiomr_amd import autotune as AT
    return AT.load(workload) is not None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_batch(wl, gb=None, threads=16):
    """The CPU checker over every view of the batch, one after another
    (one frame at a time, its thread pool inside each frame).  gb: the GPU
    batch whose framebuffers hold the batch's frames; each view is compared
    with the checker's frame (coverage, depth, colour), outside the timing."""
    import oracle as O
    from axiomr_amd import resident as H
    threads = min(threads, os.cpu_count() or 1)
    rr = O.OracleRenderer(threads)
    oscenes = {}
    dt = 0.0
    par = None
    for i, vs in enumerate(wl.views):
        sc = wl.scene(i)
        if vs.mesh not in oscenes:
            oscenes[vs.mesh] = O.OracleScene(sc)
        fb = rr.framebuffer(sc.width, sc.height)
        t0 = time.perf_counter()
        O.fb_clear(fb, H.BLACK.packed_bgra(), float("inf"))
        rr.draw(fb, oscenes[vs.mesh], H.params_of(sc))
        dt += time.perf_counter() - t0
        if gb is not None:
            ref = O.fb_read(fb, sc.width, sc.height)
            par = merge_parity(par, compare_frames(gb.fbs[i].read(), ref))
        O.fb_destroy(fb)
    for o in oscenes.values():
        o.close()
    rr.close()
    return dict(views_per_s=len(wl.views) / dt, seconds=dt, threads=threads, parity=par)


def cpu_serve(wl, n=24, threads=16):
    """Per-request CPU render time (the checker, clear + draw) over a sample
    of the served workload's request mix."""
    import oracle as O
    from axiomr_amd import resident as H
    threads = min(threads, os.cpu_count() or 1)
    rr = O.OracleRenderer(threads)
    oscenes = {}
    times = []
    for vs in wl.requests(n, seed=99):
        sc = wl.scene(vs)
        if vs.mesh not in oscenes:
            oscenes[vs.mesh] = O.OracleScene(sc)
        fb = rr.framebuffer(sc.width, sc.height)
        t0 = time.perf_counter()
        O.fb_clear(fb, H.BLACK.packed_bgra(), float("inf"))
        rr.draw(fb, oscenes[vs.mesh], H.params_of(sc))
        times.append(time.perf_counter() - t0)
        O.fb_destroy(fb)
    for o in oscenes.values():
        o.close()
    rr.close()
    t = np.array(times)
    return dict(requests=n, mean_ms=float(t.mean() * 1e3), p50_ms=float(np.percentile(t, 50) * 1e3),
                p99_ms=float(np.percentile(t, 99) * 1e3), frames_per_s=float(1.0 / t.mean()), threads=threads)


def cpu_baseline(scene, budget_s=12.0, threads=16, min_frames=1, max_frames=50):
    """The CPU checker (restated reference tiled pipeline with its thread
    pool) on the same frame, repeated until ~budget_s of CPU work (at least
    min_frames, at most max_frames frames); median frame time.  Also returns
    the last frame (bgra, depth, ids) for the parity check."""
    import oracle as O
    from axiomr_amd import resident as H
    threads = min(threads, os.cpu_count() or 1)
    osc = O.OracleScene(scene)
    rr = O.OracleRenderer(threads)
    fb = rr.framebuffer(scene.width, scene.height)
    p = H.params_of(scene)
    times = []
    t_start = time.perf_counter()
    while True:
        O.fb_clear(fb, H.BLACK.packed_bgra(), float("inf"))
        t0 = time.perf_counter()
        rr.draw(fb, osc, p)
        times.append(time.perf_counter() - t0)
        if len(times) >= max_frames or (len(times) >= min_frames and time.perf_counter() - t_start > budget_s):
            break
    phases = rr.phase_times()
    ref = O.fb_read(fb, scene.width, scene.height)
    O.fb_destroy(fb)
    rr.close()
    osc.close()
    med = float(np.median(times))
    return dict(value=1.0 / med, frames=len(times), median_s=med, threads=threads, frame=ref,
                phases_s={k: v / len(times) for k, v in phases.items()})


def measured_traffic(kname, tunables, path=None, instance=""):
    """HBM bytes per launch of `kname` from a committed --pmc summary, if it
    was taken with the same tunables (else None); `instance`: a substring of
    the template arguments that picks one instantiation ("true": the
    batch's, "false": the plain frame's without an id plane)."""
    path = TRAFFIC_FILE if path is None else path
    if not path.exists():
        return None
    data = json.loads(path.read_text())
    if data.get("tunables") != tunables:
        return None
    total = 0.0
    for part in kname.split("+"):   # "k_shade+k_finish": the sum of both launches
        hits = [v for k, v in data.get("kernels", {}).items()
                if (k.startswith(part + "<") and instance in k) or (k == part and not instance)]
        if len(hits) != 1:
            return None
        total += hits[0]["traffic_bytes"]
    return int(round(total))


def copy_calibration(mib, reps=5):
    """A streaming read+write kernel over `mib` MiB (int32 add, torch's
    elementwise kernel): under rocprofv3 --pmc its FETCH_SIZE and WRITE_SIZE
    per dispatch each correspond to exactly mib MiB."""
    import torch
    n = (mib << 20) // 4
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        torch.add(a, 1, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--tile", type=int, default=0, help="bin edge when no tuned config is used")
    ap.add_argument("--untuned", action="store_true", help="ignore axiomr_amd/tuning/mi355x.json")
    ap.add_argument("--n", type=int, default=224, help="icosphere frequency (20 n^2 faces)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-form leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 800x600 UV-sphere and 4M-triangle soup legs")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--no-serve", action="store_true", help="skip the served-form leg")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="stream length per arrival rate")
    ap.add_argument("--serve-lanes", type=int, default=3)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--raster-mode", choices=("binned", "visibility"), default=None,
                    help="override the tuned raster mode")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain-copy kernel used to scale the --pmc byte counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU design is described but not run", file=sys.stderr)
        sys.exit(2)
    global RASTER_MODE
    if args.raster_mode:
        RASTER_MODE = {"binned": 0, "visibility": 1}[args.raster_mode]
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from axiomr_amd import scenes as S
    t = time.perf_counter()
    scene = S.icosphere_scene(n=args.n)
    log(f"[bench] scene {scene.name}: {scene.nfaces} faces, {scene.vertices.shape[0]} vertices, "
        f"checksum {scene.checksum()} ({time.perf_counter() - t:.1f}s)")

    r = run_frames(scene, args.steps, args.warmup, args.tile, not args.untuned)
    fps = args.steps / r["seconds"]
    ms = 1000.0 * r["seconds"] / args.steps
    nverts, nfaces = scene.vertices.shape[0], scene.nfaces
    W, H = scene.width, scene.height
    ntex = 2
    log(f"[bench] plain path: {fps:.1f} frames/s ({ms:.4f} ms/frame, host enqueue "
        f"{1e6 * r['enqueue_seconds'] / args.steps:.1f} us/frame), kernel ms {r['stages_ms']}")

    # roofline: each kernel's algorithmic bytes over its average duration
    # (events on its dispatches in the timed frames).  The visibility-mode
    # frame's dominant kernel is k_vis (37 % of the frame's GPU time over 200
    # frames under rocprofv3's kernel trace, k_shade 35 %); the two lie within
    # a few percent of each other and swap places in a 20-frame sample, so the
    # choice is fixed.  Binned mode: the kernel with the longest average.
    kb = frame_kernel_bytes(nverts, nfaces, W, H, r["covered"], ntex)
    kernels = {}
    for name, kms_ in r["stages_ms"].items():
        nb = kb.get(name, 0)
        kernels[name] = {"avg_ms": round(kms_, 6), "bytes_per_launch": nb,
                         "achieved_gbs": round(nb / (kms_ * 1e-3) / 1e9, 2) if kms_ > 0 else None,
                         "frac": round(nb / (kms_ * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) if kms_ > 0 else None}
    dominant = "k_vis" if "k_vis" in r["stages_ms"] else max(r["stages_ms"], key=r["stages_ms"].get)
    kms = r["stages_ms"][dominant]
    kbytes = kb.get(dominant, 0)
    # k_finish4 (16-byte writes) when the width is a multiple of 4
    kname = "k_finish4" if dominant == "k_finish" and W % 4 == 0 else dominant
    achieved = kbytes / (kms * 1e-3) / 1e9
    frame_bytes = algorithmic_bytes(nverts, nfaces, W, H, r["covered"], ntex)

    cpu, parity = None, None
    extra_configs = {}
    if not args.no_cpu:
        c = cpu_baseline(scene, args.cpu_budget)
        parity = compare_frames(r["last"], c["frame"], gpu_ids=r["checked"][2])
        # the check frame with ids must be the timed frame again
        parity["check_frame_equal"] = bool(np.array_equal(r["checked"][0], r["last"][0]) and
                                           np.array_equal(r["checked"][1].view(np.uint32),
                                                          r["last"][1].view(np.uint32)))
        log(f"[bench] parity of the last timed frame against the checker: {parity}")
        log(f"[bench] cpu baseline: {c['value']:.3f} frames/s over {c['frames']} frames, phases {c['phases_s']}")
        c1 = cpu_baseline(scene, args.cpu_budget / 2, threads=1)
        log(f"[bench] cpu baseline, 1 thread: {c1['value']:.3f} frames/s over {c1['frames']} frames")
        cpu = {"value": round(c["value"], 4), "unit": "frames/s", "cores": c["threads"], "kind": "port",
               "sample": f"{c['frames']} frames of the same 1080p 1M-triangle scene, median",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "single_thread_value": round(c1["value"], 4), "single_thread_frames": c1["frames"],
               "phases_s": {k: round(v, 5) for k, v in c["phases_s"].items()}}

    if not args.no_configs:
        # the reference's own CPU case (800x600 UV sphere, Gouraud x checker)
        # and the overdraw/tie stress (4M-triangle soup): GPU frames/s and
        # Gtri/s, the checker's frames/s, parity of the last timed frame
        for key, mk, min_f, max_f in (("uv_sphere_800x600", lambda: S.uv_sphere_scene(), 50, 200),
                                      ("soup_4M_1080p", lambda: S.soup_scene(), 5, 5)):
            t = time.perf_counter()
            sc = mk()
            log(f"[bench] scene {sc.name}: {sc.nfaces} faces ({time.perf_counter() - t:.1f}s)")
            # its own tuned tunables when the tuner has an entry for it
            rr_ = run_frames(sc, args.steps, args.warmup, args.tile, not args.untuned,
                             workload=key if _has_tuned(key) else "plain")
            gfps = args.steps / rr_["seconds"]
            rec = {"workload": sc.name, "triangles": sc.nfaces, "width": sc.width, "height": sc.height,
                   "frames_per_s": round(gfps, 2), "gtri_per_s": round(gfps * sc.nfaces / 1e9, 4),
                   "ms_per_frame": round(1000.0 * rr_["seconds"] / args.steps, 5),
                   "kernel_ms": {k: round(v, 6) for k, v in rr_["stages_ms"].items()},
                   "covered_pixels": rr_["covered"], "tunables": rr_["config"]}
            if not args.no_cpu:
                cc = cpu_baseline(sc, args.cpu_budget, min_frames=min_f, max_frames=max_f)
                rec["cpu_frames_per_s"] = round(cc["value"], 4)
                rec["cpu_frames"] = cc["frames"]
                rec["cpu_threads"] = cc["threads"]
                rec["cpu_phases_s"] = {k: round(v, 5) for k, v in cc["phases_s"].items()}
                rec["parity"] = compare_frames(rr_["last"], cc["frame"], gpu_ids=rr_["checked"][2])
            log(f"[bench] {key}: {rec}")
            extra_configs[key] = rec

    batch = None
    if not args.no_batch:
        t = time.perf_counter()
        wl = S.batch_workload()
        log(f"[bench] batch workload: {len(wl.views)} views, {wl.total_faces} faces, {wl.total_pixels} pixels "
            f"({time.perf_counter() - t:.1f}s)")
        b = run_batch(wl, args.batch_reps, args.tile, not args.untuned)
        cb = None if args.no_cpu else cpu_batch(wl, b["gb"])
        b["gb"].close()
        b["ctx"].close()
        if cb:
            log(f"[bench] batch on the CPU checker: {cb['views_per_s']:.1f} views/s ({cb['threads']} threads), "
                f"parity {cb['parity']}")
        log(f"[bench] batch: plain {b['plain_views_per_s']:.0f} views/s (queued), {b['sync_views_per_s']:.0f} views/s "
            f"(synchronous drawMesh), batched {b['batch_views_per_s']:.0f} views/s, x{b['speedup']:.2f} / "
            f"x{b['speedup_vs_sync']:.2f}; kernel ms {b['stages_ms']}")
        bkb = batch_kernel_bytes(wl, b["covered"], ntex)
        bdom = max(b["stages_ms"], key=b["stages_ms"].get)
        bms = b["stages_ms"][bdom]
        bach = bkb.get(bdom, 0) / (bms * 1e-3) / 1e9 if bms > 0 else 0.0
        batch = {"workload": f"{len(wl.views)} views of 32 displaced icospheres (5k-500k triangles, log-uniform), "
                             "256x256..1920x1080 log-uniform, Phong + textures, one graph launch per batch",
                 "views": b["views"], "triangles": b["faces"], "pixels": b["pixels"], "covered_pixels": b["covered"],
                 "batched_views_per_s": round(b["batch_views_per_s"], 1),
                 "plain_views_per_s": round(b["plain_views_per_s"], 1),
                 "speedup_vs_plain": round(b["speedup"], 3),
                 "sync_views_per_s": round(b["sync_views_per_s"], 1),
                 "speedup_vs_sync": round(b["speedup_vs_sync"], 3),
                 "batched_gtri_per_s": round(b["faces"] / (b["batch_ms"] * 1e-3) / 1e9, 3),
                 "batch_ms": round(b["batch_ms"], 4), "kernel_ms": {k: round(v, 5) for k, v in b["stages_ms"].items()},
                 "roofline": {"bound": "hbm", "kernel": bdom, "achieved": round(bach, 2), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(bach / HBM_PEAK_GBS, 5),
                              "bytes_per_launch": bkb.get(bdom, 0), "avg_launch_ms": round(bms, 5),
                              "traffic": measured_traffic(bdom, b["config"], BATCH_TRAFFIC_FILE, "true")},
                 "batch_algorithmic_bytes": sum(bkb.values()),
                 "batch_roofline_frac": round(sum(bkb.values()) / (b["batch_ms"] * 1e-3) / (HBM_PEAK_GBS * 1e9), 5),
                 "tunables": b["config"], "plain_tunables": b["plain_config"],
                 "cpu_views_per_s": round(cb["views_per_s"], 2) if cb else None,
                 "cpu_threads": cb["threads"] if cb else None,
                 "parity": cb["parity"] if cb else None}

    served = None
    if not args.no_serve:
        t = time.perf_counter()
        swl = S.serve_workload()
        log(f"[bench] serve workload: {len(swl.meshes)} meshes, {swl.total_faces} faces "
            f"({time.perf_counter() - t:.1f}s)")
        budget = int(args.serve_budget_gb * (1 << 30))
        sv = run_serve(swl, SERVE_RATES, args.serve_seconds, args.serve_lanes, budget, args.tile, not args.untuned)
        cs = None if args.no_cpu else cpu_serve(swl)
        if cs:
            log(f"[bench] serve requests on the CPU checker: mean {cs['mean_ms']:.1f} ms, p99 {cs['p99_ms']:.1f} ms")
        log(f"[bench] serve capacity {sv['capacity_frames_per_s']:.0f} frames/s, plain "
            f"{sv['plain_frames_per_s']:.0f} frames/s, x{sv['speedup']:.2f}")
        st = sv["stats"]
        served = {"workload": f"Poisson request stream, {len(swl.meshes)} resident displaced icospheres "
                              "(10k-2M triangles, log-uniform), 50% 512^2 / 35% 1280x720 / 15% 1920x1080, "
                              "random orbit cameras, Phong + textures, 3 point lights",
                  "lanes": args.serve_lanes, "hbm_budget_bytes": budget, "rates": sv["rates"],
                  "capacity_frames_per_s": round(sv["capacity_frames_per_s"], 1),
                  "capacity_gtri_per_s": round(sv["capacity_gtri_per_s"], 3),
                  "plain_frames_per_s": round(sv["plain_frames_per_s"], 1),
                  "speedup_vs_plain": round(sv["speedup"], 3),
                  "fb_peak_bytes": st["fb_peak_bytes"], "fixed_bytes": st["fixed_bytes"],
                  "batches": st["batches"], "max_views_in_batch": st["max_views_in_batch"],
                  "max_lanes_busy": st["max_lanes_busy"], "overflow_retries": st["overflow_retries"],
                  "tunables": sv["config"],
                  "cpu_request_ms": ({"mean": round(cs["mean_ms"], 2), "p50": round(cs["p50_ms"], 2),
                                      "p99": round(cs["p99_ms"], 2), "requests": cs["requests"],
                                      "threads": cs["threads"]} if cs else None)}

    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain per-frame GPU path, 1920x1080, displaced icosphere n={args.n} "
                               f"({nfaces} triangles), Phong + 2048^2 albedo + 2048^2 normal map (bilinear), "
                               f"3 point lights, FP32 depth, BGRA8",
                   "tunables": r["config"], "covered_pixels": r["covered"]},
        "gtri_per_s": round(fps * nfaces / 1e9, 4),
        "host_enqueue_us_per_frame": round(1e6 * r["enqueue_seconds"] / args.steps, 2),
        "frame_algorithmic_bytes": frame_bytes,
        "frame_roofline_frac": round(frame_bytes * fps / (HBM_PEAK_GBS * 1e9), 5),
        "kernels": kernels,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(kname, r["config"], instance="false"),
                     "bytes_per_launch": kbytes, "avg_launch_ms": round(kms, 6)},
        "cpu_baseline": cpu,
        "parity": parity,
        "configs": extra_configs or None,
        "batched": batch,
        "served": served,
    }
    print(json.dumps(out), flush=True)
    # a fast frame whose pixels differ from the reference's is not a result
    bad = []
    if parity is not None and not (parity_ok(parity) and parity["check_frame_equal"]):
        bad.append(f"1080p frame {parity}")
    for key, rec in extra_configs.items():
        if rec.get("parity") is not None and not parity_ok(rec["parity"]):
            bad.append(f"{key} {rec['parity']}")
    if batch and batch.get("parity") is not None and not parity_ok(batch["parity"]):
        bad.append(f"batch {batch['parity']}")
    if bad:
        log("[bench] PARITY FAILED: " + "; ".join(bad))
        sys.exit(1)


if __name__ == "__main__":
    main()
