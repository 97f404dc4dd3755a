"""Passes the host leaves out: a draw whose mesh bounds show that no face
crosses a frustum plane does not launch the clipper, and one whose faces
cannot cover more than the one-lane footprint does not launch the
large-triangle pass.  The decision must never drop work: every scene here
equals the checker, and no draw reports a skipped pass that had work
(skip_faults), whether or not its passes were skipped."""
import numpy as np
import pytest

import harness as H
from axiomr_amd import _native as N
from axiomr_amd import scenes as S

pytestmark = pytest.mark.gpu


def _draw_and_count(ctx, scene, flags=0):
    ctx.set_profiling(True)         # resets the counters
    got = H.render_gpu(ctx, scene)
    st = ctx.stats()
    ctx.set_profiling(False)
    return got, st.passes_skipped, st.skip_faults


@pytest.mark.parametrize("case", ["far", "near_plane", "inside", "low_poly_1080p", "soup"])
def test_skips_keep_every_pixel(ctx, case):
    ctx.set_config(small_footprint=128)
    if case == "far":           # the whole mesh well inside the frustum, small triangles
        sc = S.icosphere_scene(n=128, width=640, height=480, tex=64)
    elif case == "near_plane":  # the camera close enough that faces cross the near plane
        sc = S.icosphere_scene(n=24, width=320, height=240, tex=64, radius=1.05)
    elif case == "inside":      # the camera inside the sphere
        sc = S.icosphere_scene(n=24, width=320, height=240, tex=64, radius=0.5)
    elif case == "low_poly_1080p":   # large triangles
        sc = S.icosphere_scene(n=3, width=1920, height=1080, tex=64)
    else:                       # depths out to the far plane: clipping
        sc = S.soup_scene(ntri=20_000, width=320, height=240, seed=4)
    got, skipped, faults = _draw_and_count(ctx, sc)
    H.assert_parity(H.compare(got, H.render_oracle(sc)), case)
    assert faults == 0, f"{case}: a skipped pass had work"
    mode = ctx.config()["raster_mode"]
    if case == "far":
        # clip skipped, and in visibility mode the large-triangle pass too
        assert skipped == (2 if mode == N.RASTER_VISIBILITY else 1), (case, skipped)
    if case in ("near_plane", "inside", "soup"):
        assert skipped == 0, (case, skipped)


@pytest.mark.parametrize("crossing", [False, True])
def test_batch_of_inside_views_equals_solo(ctx, crossing):
    """A binned batch whose every view the bounds put inside the frustum runs
    its setup without the classification and launches no clipper; one view
    that crosses the near plane brings the clipper back for the batch and the
    classification for its own faces.  Either way every view equals its solo
    draw, and the crossing view equals the checker."""
    import dataclasses
    wl = S.batch_workload(nviews=10, nmeshes=4, fmin=500, fmax=20_000, wmin=64, wmax=320, hmin=48, hmax=240,
                          tex=64, seed=11)
    for i, v in enumerate(wl.views):
        p = np.asarray(v.camera_pos, np.float64)
        d = 1.05 if (crossing and i == 3) else 4.0   # 1.05: on the displaced surface, faces cross the near plane
        wl.views[i] = dataclasses.replace(v, camera_pos=tuple(float(x) for x in p * (d / np.linalg.norm(p))))
    gb = H.GpuBatch(ctx, wl)
    try:
        ctx.set_profiling(True)         # resets the counters; kernels timed per stage
        gb.batch.launch()
        ctx.sync()
        st = ctx.stats()
        ctx.set_profiling(False)
        assert st.skip_faults == 0
        clip_launched = st.stage_ms_total[1] > 0     # stage 1: k_clip in both modes
        if ctx.config()["raster_mode"] != N.RASTER_VISIBILITY:
            assert clip_launched == crossing, (crossing, list(st.stage_ms_total))
        gb.batch.launch()               # the captured graph
        ctx.sync()
        batched = [fb.read() for fb in gb.fbs]
        for i in range(len(wl.views)):
            gb.draw_solo(i)
            solo = gb.fbs[i].read()
            for a, b in zip(batched[i], solo):
                assert (a.view(np.uint8) == b.view(np.uint8)).all(), f"view {i} differs from its solo render"
        if crossing:
            H.assert_parity(H.compare(batched[3], H.render_oracle(wl.scene(3))), "crossing view")
    finally:
        gb.close()
