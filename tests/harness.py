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


def shader_of(scene):
    s = dict(scene.shader)
    kind = s.pop("kind")
    cls = {0: R.FlatShader, 1: R.PhongShader, 2: R.PBRShader, 3: R.GouraudShader, 4: R.PhongPointShader,
           5: R.FlatColorShader}[kind]
    return cls(**s)


def camera_of(scene):
    cam = R.Camera(scene.camera_pos, scene.camera_target, scene.fov, scene.width / scene.height)
    cam.setViewport(0, 0, scene.width, scene.height)
    return cam


def params_of(scene, model=None, flags=0):
    return R.draw_params(scene.model if model is None else model, camera_of(scene), shader_of(scene), flags)


class GpuScene:
    """Device-resident copies of a SceneData's textures and mesh."""

    def __init__(self, ctx, scene):
        self.ctx = ctx
        self.textures = {k: R.Texture(ctx, t.rgba, t.filter) for k, t in scene.textures.items()}
        mats = []
        for m in scene.materials:
            g = lambda n: self.textures[n] if n else None  # noqa: E731
            mats.append(R.Material(g(m.diffuse), g(m.metallic), g(m.bump), g(m.roughness), g(m.ao),
                                   m.specular_exponent))
        groups = [R.MaterialGroup(*g) for g in scene.groups]
        self.mesh = R.Mesh(ctx, scene.vertices, scene.faces, groups, mats, scene.face_colors)

    def close(self):
        self.mesh.close()
        for t in self.textures.values():
            t.close()


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


class GpuBatch:
    """A BatchWorkload made resident: one texture set, one mesh per workload
    mesh, one framebuffer per view, and the batch handle."""

    def __init__(self, ctx, wl, with_ids=True, tile=None):
        if tile:
            ctx.set_tile_size(tile)
        self.ctx = ctx
        self.wl = wl
        self.textures = {k: R.Texture(ctx, t.rgba, t.filter) for k, t in wl.textures.items()}
        m = wl.material
        g = lambda n: self.textures[n] if n else None  # noqa: E731
        mat = R.Material(g(m.diffuse), g(m.metallic), g(m.bump), g(m.roughness), g(m.ao), m.specular_exponent)
        self.meshes = [R.Mesh(ctx, v, f, [R.MaterialGroup(0, 0, f.shape[0])], [mat]) for v, f in wl.meshes]
        self.fbs = [R.Framebuffer(ctx, vs.width, vs.height, with_ids=with_ids) for vs in wl.views]
        self.params = [params_of(wl.scene(i)) for i in range(len(wl.views))]
        self.batch = R.Batch(ctx, self.fbs, [self.meshes[vs.mesh] for vs in wl.views], self.params)
        self.batch.set_clear(BLACK)

    def draw_solo(self, i, flags=0):
        """View i through the plain per-frame path (clear + drawMesh)."""
        import ctypes as C
        fb = self.fbs[i]
        fb.clear(BLACK)
        p = self.params[i]
        p.flags = flags
        self.ctx.check(self.ctx._lib.axr_draw(self.ctx.handle, fb.handle, self.meshes[self.wl.views[i].mesh].handle,
                                              C.byref(p)))
        p.flags = 0

    def close(self):
        self.batch.close()
        for fb in self.fbs:
            fb.close()
        for m in self.meshes:
            m.close()
        for t in self.textures.values():
            t.close()
