"""Tune the binning pipeline for the benchmark workloads on this GPU and
store the result in axiomr_amd/tuning/mi355x.json:

    python -m axiomr_amd.tune_workloads [plain] [batch]

plain: one 1920x1080 frame of the 1,003,520-triangle icosphere;
batch: the 256-view / 32-mesh batch in one graph launch.
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
    scene = S.icosphere_scene(n=224)
    gs = RS.GpuScene(ctx, scene)
    fb = Framebuffer(ctx, scene.width, scene.height)
    p = RS.params_of(scene, flags=1)
    L = ctx._lib

    def make_frame():
        def frame():
            L.axr_fb_clear(fb.handle, RS.BLACK.packed_bgra(), float("inf"))
            L.axr_draw(ctx.handle, fb.handle, gs.mesh.handle, C.byref(p))
        return frame

    res = AT.tune(ctx, make_frame, log, reps=5, inner=20)
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


def main(argv):
    which = argv or ["plain", "batch"]
    ctx = Context(0)
    for w in which:
        res = {"plain": tune_plain, "batch": tune_batch}[w](ctx)
        log(f"[autotune] {w}: best {res['best']} {res['seconds'] * 1e3:.3f} ms")
        AT.store(w, res)
        print(f"{w} {res['best']} {res['seconds'] * 1e3:.4f} ms", flush=True)
    ctx.close()


if __name__ == "__main__":
    main(sys.argv[1:])
