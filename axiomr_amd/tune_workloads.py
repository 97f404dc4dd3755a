"""Tune the binning pipeline for the benchmark workloads on this GPU and
store the result in axiomr_amd/tuning/mi355x.json:

    python -m axiomr_amd.tune_workloads [plain] [batch] [serve] [uv_sphere_800x600] [soup_4M_1080p]

plain: one 1920x1080 frame of the 1,003,520-triangle icosphere;
uv_sphere_800x600: one frame of the 20,000-triangle UV sphere (Gouraud x
checker, the reference application's scene);
soup_4M_1080p: one 1920x1080 frame of the 4M-triangle overdraw soup;
batch: the 256-view / 32-mesh batch in one graph launch;
serve: a burst of 400 requests of the served workload (64 meshes, mixed
resolutions) through a 3-lane render server.
"""
from __future__ import annotations

import ctypes as C
import sys

from . import autotune as AT
from . import resident as RS
from . import scenes as S
from .render import Context, Framebuffer


def log(m):
    print(m, file=sys.stderr, flush=True)


def tune_plain(ctx):
    return tune_scene(ctx, S.icosphere_scene(n=224), "plain", 200)


def tune_scene(ctx, scene, workload, inner):
    """Asynchronous clear + draw frames of one scene into one framebuffer,
    `inner` of them per timed run."""
    gs = RS.GpuScene(ctx, scene)
    fb = Framebuffer(ctx, scene.width, scene.height)
    p = RS.params_of(scene, flags=1)
    L = ctx._lib

    def make_frame():
        def frame():
            L.axr_fb_clear(fb.handle, RS.BLACK.packed_bgra(), float("inf"))
            L.axr_draw(ctx.handle, fb.handle, gs.mesh.handle, C.byref(p))
        return frame

    cached = AT.load(workload) or AT.load("plain")
    if cached:
        ctx.set_config(**cached["config"])
    res = AT.tune(ctx, make_frame, log, reps=5, inner=inner)
    fb.close()
    gs.close()
    return res


def tune_batch(ctx):
    wl = S.batch_workload()
    gb = RS.GpuBatch(ctx, wl, with_ids=False)

    def make_frame():
        gb.rebuild_batch()
        return gb.batch.launch

    res = AT.tune(ctx, make_frame, log, reps=3, inner=2)
    gb.close()
    return res


def tune_serve(ctx):
    import numpy as np
    from .serve import Server
    wl = S.serve_workload()
    gs = RS.GpuServeScene(ctx, wl)
    reqs = [gs.request(v) for v in wl.requests(400, seed=11)]
    zeros = np.zeros(len(reqs))
    holder = {}

    def make_frame():
        if "srv" in holder:
            holder["srv"].close()
        holder["srv"] = srv = Server(ctx, lanes=3, max_batch_views=64)
        return lambda: srv.replay(reqs, zeros)

    res = AT.tune(ctx, make_frame, log, reps=3, inner=1)
    holder["srv"].close()
    gs.close()
    return res


def main(argv):
    which = argv or ["plain", "batch", "serve", "uv_sphere_800x600", "soup_4M_1080p"]
    runs = {"plain": tune_plain, "batch": tune_batch, "serve": tune_serve,
            "uv_sphere_800x600": lambda c: tune_scene(c, S.uv_sphere_scene(), "uv_sphere_800x600", 200),
            "soup_4M_1080p": lambda c: tune_scene(c, S.soup_scene(), "soup_4M_1080p", 10)}
    ctx = Context(0)
    for w in which:
        res = runs[w](ctx)
        log(f"[autotune] {w}: best {res['best']} {res['seconds'] * 1e3:.3f} ms")
        AT.store(w, res)
        print(f"{w} {res['best']} {res['seconds'] * 1e3:.4f} ms", flush=True)
    ctx.close()


if __name__ == "__main__":
    main(sys.argv[1:])
