"""Make SceneData / BatchWorkload resident on the GPU (textures, meshes,
framebuffers, batch handles) and build the draw uniforms the way the
reference's drawMesh receives them (camera view-projection + shader
fields)."""
from __future__ import annotations

import ctypes as C

from . import render as R

BLACK = R.Color(0, 0, 0, 255)


def shader_of(scene):
    s = dict(scene.shader)
    kind = s.pop("kind")
    cls = {0: R.FlatShader, 1: R.PhongShader, 2: R.PBRShader, 3: R.GouraudShader, 4: R.PhongPointShader,
           5: R.FlatColorShader}[kind]
    return cls(**s)


def camera_of(scene):
    cam = R.Camera(scene.camera_pos, scene.camera_target, scene.fov, scene.width / scene.height)
    cam.setViewport(0, 0, scene.width, scene.height)
    cam.update(0.0)
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

    def rebuild_batch(self):
        """Re-create the batch handle (a batch binds the context's tunables
        when it is created)."""
        self.batch.close()
        self.batch = R.Batch(self.ctx, self.fbs, [self.meshes[vs.mesh] for vs in self.wl.views], self.params)
        self.batch.set_clear(BLACK)


class GpuServeScene:
    """A ServeWorkload made resident: one texture set, one material, one
    mesh per workload mesh.  `request(vs)` builds the server request for a
    ViewSpec."""

    def __init__(self, ctx, wl):
        from .serve import request
        self._request = request
        self.ctx = ctx
        self.wl = wl
        self.textures = {k: R.Texture(ctx, t.rgba, t.filter) for k, t in wl.textures.items()}
        m = wl.material
        g = lambda n: self.textures[n] if n else None  # noqa: E731
        mat = R.Material(g(m.diffuse), g(m.metallic), g(m.bump), g(m.roughness), g(m.ao), m.specular_exponent)
        self.meshes = [R.Mesh(ctx, v, f, [R.MaterialGroup(0, 0, f.shape[0])], [mat]) for v, f in wl.meshes]

    def params(self, vs):
        return params_of(self.wl.scene(vs))

    def request(self, vs):
        return self._request(self.meshes[vs.mesh], vs.width, vs.height, self.params(vs))

    def draw_plain(self, vs, with_ids=False):
        """The same request through the plain per-frame path (clear +
        drawMesh into a new framebuffer); returns (bgra, depth, ids)."""
        fb = R.Framebuffer(self.ctx, vs.width, vs.height, with_ids=with_ids)
        try:
            fb.clear(BLACK)
            p = self.params(vs)
            self.ctx.check(self.ctx._lib.axr_draw(self.ctx.handle, fb.handle, self.meshes[vs.mesh].handle,
                                                  C.byref(p)))
            return fb.read()
        finally:
            fb.close()

    def close(self):
        for m in self.meshes:
            m.close()
        for t in self.textures.values():
            t.close()
