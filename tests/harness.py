"""Shared helpers of the test suite: run one SceneData through the GPU path
(libaxiomr_amd.so) and through the CPU checker (oracle/), and compare."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import axiomr_amd as axr  # noqa: E402
from axiomr_amd import render as R  # noqa: E402

BLACK = R.Color(0, 0, 0, 255)
EMPTY_ID = 0xFFFFFFFF


from axiomr_amd.resident import GpuBatch, GpuScene, camera_of, params_of, shader_of  # noqa: E402,F401


def render_gpu(ctx, scene, draws=None, tile=None, clear=BLACK, with_ids=True):
    """draws: list of (model, params) or None for one draw of scene.model."""
    if tile:
        ctx.set_tile_size(tile)
    gs = GpuScene(ctx, scene)
    fb = R.Framebuffer(ctx, scene.width, scene.height, with_ids=with_ids)
    try:
        fb.clear(clear)
        for p in (draws or [params_of(scene)]):
            ctx.check(ctx._lib.axr_draw(ctx.handle, fb.handle, gs.mesh.handle, R.C.byref(p)))
        return fb.read()
    finally:
        fb.close()
        gs.close()


def render_oracle(scene, draws=None, threads=0, clear=BLACK):
    import oracle as O
    os_ = O.OracleScene(scene)
    rr = O.OracleRenderer(threads)
    fb = rr.framebuffer(scene.width, scene.height)
    try:
        O.fb_clear(fb, clear.packed_bgra(), float("inf"))
        for p in (draws or [params_of(scene)]):
            rr.draw(fb, os_, p)
        return O.fb_read(fb, scene.width, scene.height)
    finally:
        O.fb_destroy(fb)
        rr.close()
        os_.close()


def compare(gpu, ref):
    """Parity report: coverage/depth/winner must match bit for bit, colour to
    within 1 LSB per channel."""
    gc, gd, gi = gpu
    rc, rd, ri = ref
    cov_g = gd.view(np.uint32) != np.float32(np.inf).view(np.uint32)
    cov_r = rd.view(np.uint32) != np.float32(np.inf).view(np.uint32)
    diff = np.abs(gc.astype(np.int16) - rc.astype(np.int16))
    rep = {
        "pixels": int(gd.size),
        "covered": int(cov_r.sum()),
        "coverage_mismatch": int((cov_g != cov_r).sum()),
        "depth_mismatch": int((gd.view(np.uint32) != rd.view(np.uint32)).sum()),
        "id_mismatch": int((gi != ri).sum()) if gi is not None and ri is not None else 0,
        "color_max": int(diff.max()) if diff.size else 0,
        "color_mean": float(diff.mean()) if diff.size else 0.0,
        "color_over_1": int((diff > 1).sum()),
    }
    return rep


def assert_parity(rep, tag=""):
    ok = (rep["coverage_mismatch"] == 0 and rep["depth_mismatch"] == 0 and rep["id_mismatch"] == 0
          and rep["color_max"] <= 1)
    assert ok, f"parity failed {tag}: {rep}"
