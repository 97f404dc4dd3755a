"""A camera path (the reference's fly camera replayed over recorded input)
drawn as one batch: every view matches the CPU checker given the same
uniforms, and the batch equals drawing the path frame by frame through the
pipeline interface."""
import numpy as np
import pytest

import harness as H
from axiomr_amd import render as R
from axiomr_amd import scenes as S

pytestmark = pytest.mark.gpu


def recorded_flight(n):
    """Frame i: drag right/down while the right button is held (released every
    7th frame), W held on most frames, A on some."""
    ins = []
    for i in range(n):
        keys = ("W" if i % 4 else "") + ("A" if i % 5 == 2 else "")
        ins.append(R.InputState((300.0 + 6.0 * i, 200.0 + 2.5 * i), i % 7 != 0, keys))
    return ins, np.full(n, 1.0 / 30.0, np.float32)


def test_camera_path_as_one_batch(ctx):
    sc = S.icosphere_scene(n=24, width=256, height=192, tex=128)
    gs = H.GpuScene(ctx, sc)
    cam = R.Camera(sc.camera_pos, sc.camera_target, sc.fov, sc.width / sc.height)
    cam.setViewport(0, 0, sc.width, sc.height)
    ins, dts = recorded_flight(12)
    vp, pos = cam.path(ins, dts)
    assert len({bytes(v) for v in vp}) == 12            # every frame moved or turned
    base = H.params_of(sc)
    params = []
    for i in range(12):
        p = type(base).from_buffer_copy(base)
        p.view_proj[:] = [float(x) for x in vp[i]]
        p.camera_pos[:] = [float(x) for x in pos[i]]
        params.append(p)
    fbs = [R.Framebuffer(ctx, sc.width, sc.height, with_ids=True) for _ in range(12)]
    batch = R.Batch(ctx, fbs, [gs.mesh] * 12, params)
    try:
        batch.set_clear(H.BLACK)
        batch.launch()
        ctx.sync()
        covered = 0
        for i, fb in enumerate(fbs):
            got = fb.read()
            rep = H.compare(got, H.render_oracle(sc, draws=[params[i]]))
            H.assert_parity(rep, f"camera path frame {i}")
            covered += rep["covered"]
        assert covered > 12 * 3000
        # the same path frame by frame through the pipeline interface
        live = R.Camera(sc.camera_pos, sc.camera_target, sc.fov, sc.width / sc.height)
        live.setViewport(0, 0, sc.width, sc.height)
        fb = R.Framebuffer(ctx, sc.width, sc.height, with_ids=True)
        pipe = R.TiledPipeline(live, fb)
        pipe.setShader(H.shader_of(sc))
        try:
            for i in range(12):
                live.handleInput(ins[i])
                live.update(float(dts[i]))
                if i not in (0, 5, 11):
                    continue
                fb.clearColor(H.BLACK)
                fb.clearDepth()
                pipe.drawMesh(sc.model, gs.mesh)
                for a, b in zip(fb.read(), fbs[i].read()):
                    assert (a.view(np.uint8) == b.view(np.uint8)).all(), f"frame {i}"
        finally:
            fb.close()
    finally:
        batch.close()
        for fb in fbs:
            fb.close()
        gs.close()
