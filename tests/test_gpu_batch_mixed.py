"""A batch whose views differ in every way the batched kernels branch on:
all six shader kinds in one launch (the raster and resolve kernels built for
any shader, and vertex-pass jobs of several kinds), a mesh whose material
groups are listed out of face order (per-face materials), views whose faces
cross the near plane and the frustum sides (the clipper, with the view of
each face found by search), a triangle soup with exact depth ties, and a
rotated, scaled model for the flat shader's normal matrix.  Every view must
be bit for bit the same as the same view drawn alone, with and without the
batch's clear, and match the CPU checker."""
import numpy as np
import pytest

import harness as H
from axiomr_amd import render as R
from axiomr_amd import scenes as S

pytestmark = pytest.mark.gpu


def mixed_scenes():
    base = S.icosphere_scene(n=32, width=200, height=150, tex=128, yaw=-40.0)
    out = []
    out.append(S.with_shader(base, {"kind": 0, "lightDirection": (0.2, -1.0, -0.3)}, name="flat"))
    out.append(S.with_shader(base, {"kind": 1, "lightDirection": (0.2, -1.0, -0.3), "lightColor": (0.6, 0.6, 0.6)},
                             name="phong"))
    out.append(S.pbr_scene(n=24, width=160, height=120, tex=64))
    out.append(S.with_shader(base, {"kind": 3, "lightDirection": (0.2, -1.0, -0.3)}, name="gouraud"))
    out.append(S.icosphere_scene(n=32, width=176, height=144, tex=128, yaw=60.0))          # point lights
    out.append(S.soup_scene(ntri=20_000, width=240, height=136, seed=4))                    # flat colour, ties
    groups = S.icosphere_scene(n=32, width=208, height=160, tex=128, yaw=15.0)
    nf = groups.nfaces
    a, b = nf // 3, 2 * nf // 3
    groups.groups = [(1, b, nf - b), (0, 0, a), (1, a, b - a)]
    groups.materials = [groups.materials[0], S.MaterialData(diffuse="normal", bump="albedo", specular_exponent=0.3)]
    groups.name = "groups_out_of_order"
    out.append(groups)
    out.append(S.icosphere_scene(n=48, width=192, height=144, tex=128, radius=1.12, yaw=10.0, pitch=5.0))  # clipped
    rot = S.with_shader(S.uv_sphere_scene(width=160, height=120), {"kind": 0, "lightDirection": (0.5, -0.7, -0.2)},
                        name="flat_rotated")
    c, s_ = np.cos(0.7), np.sin(0.7)
    rot.model = np.array([[c, 0, -s_, 0], [0, 1.3, 0, 0], [s_, 0, c, 0], [0.1, -0.2, -0.3, 1]], np.float32).reshape(16)
    out.append(rot)
    return out


def test_mixed_batch_equals_solo_and_checker(ctx):
    scenes = mixed_scenes()
    print(f"mixed batch, raster mode {ctx.config()['raster_mode']}: {[s.name for s in scenes]}", flush=True)
    gss = [H.GpuScene(ctx, sc) for sc in scenes]
    fbs = [R.Framebuffer(ctx, sc.width, sc.height, with_ids=True) for sc in scenes]
    params = [H.params_of(sc) for sc in scenes]
    batch = None
    try:
        batch = R.Batch(ctx, fbs, [g.mesh for g in gss], params)
        batch.set_clear(H.BLACK)
        batch.launch()
        ctx.sync()
        batched = [fb.read() for fb in fbs]
        # without the batch's clear the views draw over what the framebuffers
        # hold (here: the previous launch's frame), as plain draws do
        batch.set_clear(None)
        batch.launch()
        ctx.sync()
        over = [fb.read() for fb in fbs]
        for i, sc in enumerate(scenes):
            solo = H.render_gpu(ctx, sc)
            for a, b in zip(batched[i], solo):
                assert (a.view(np.uint8) == b.view(np.uint8)).all(), f"{sc.name}: batch differs from solo"
            ref = H.render_oracle(sc)
            H.assert_parity(H.compare(batched[i], ref), f"batch view {sc.name}")
            # a second draw of the same view over itself keeps every pixel:
            # the merge replaces a pixel only where the new depth is less
            for a, b in zip(over[i], batched[i]):
                assert (a.view(np.uint8) == b.view(np.uint8)).all(), f"{sc.name}: redraw without clear changed it"
        assert any(int(np.isfinite(b[1]).sum()) > 1000 for b in batched)
    finally:
        if batch is not None:
            batch.close()
        for fb in fbs:
            fb.close()
        for g in gss:
            g.close()
