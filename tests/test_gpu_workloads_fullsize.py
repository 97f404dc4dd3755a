"""The batched and served workloads at their benchmark sizes.

Batched (256 views of 32 displaced icospheres, 5k-500k triangles, 256x256
to 1920x1080): every view of one graph launch against the CPU checker
(coverage, depth and winner bit for bit, colour within 1 LSB) and against
the same view drawn alone by the plain per-frame path.

Served (64 resident icospheres of 10k-2M triangles, 50% 512^2 / 35% 720p /
15% 1080p): a stream of requests admitted together and drawn in shared
batches on three lanes; every request equals the same request drawn alone,
and a sample that includes the largest mesh at 1920x1080 equals the
checker's frame."""
import numpy as np
import pytest

import harness as H
from axiomr_amd import scenes as S
from axiomr_amd.resident import GpuServeScene
from axiomr_amd.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_wl():
    return S.batch_workload()


def test_batch_fullsize_matches_checker_and_solo(ctx, batch_wl):
    wl = batch_wl
    gb = H.GpuBatch(ctx, wl, with_ids=True)
    try:
        gb.batch.launch()
        ctx.sync()
        worst, covered = 0, 0
        for i in range(len(wl.views)):
            got = gb.fbs[i].read()
            rep = H.compare(got, H.render_oracle(wl.scene(i)))
            H.assert_parity(rep, f"batch view {i} ({wl.views[i].width}x{wl.views[i].height})")
            worst = max(worst, rep["color_max"])
            covered += rep["covered"]
            gb.draw_solo(i)
            solo = gb.fbs[i].read()
            for a, b, name in zip(got, solo, ("colour", "depth", "ids")):
                assert (a.view(np.uint8) == b.view(np.uint8)).all(), f"view {i}: {name} differs from its solo draw"
        assert covered > 20_000_000, covered
        print(f"batch: {len(wl.views)} views, {covered} covered pixels, colour max {worst}")
    finally:
        gb.close()


def test_batch_rejects_a_framebuffer_in_two_views(ctx):
    from axiomr_amd import render as R
    from axiomr_amd._native import AxrError, AXR_ERR_INVALID
    wl = S.batch_workload(nviews=3, nmeshes=2, fmin=500, fmax=2_000, wmin=32, wmax=64, hmin=32, hmax=64, tex=16)
    gb = H.GpuBatch(ctx, wl)
    try:
        fbs = [gb.fbs[0], gb.fbs[1], gb.fbs[0]]
        with pytest.raises(AxrError) as e:
            R.Batch(ctx, fbs, [gb.meshes[v.mesh] for v in wl.views], gb.params)
        assert e.value.code == AXR_ERR_INVALID
        # a framebuffer or mesh a live batch draws cannot be destroyed
        assert ctx._lib.axr_fb_destroy(gb.fbs[1].handle) == AXR_ERR_INVALID
        assert ctx._lib.axr_mesh_destroy(gb.meshes[wl.views[1].mesh].handle) == AXR_ERR_INVALID
        gb.batch.launch()
        ctx.sync()
    finally:
        gb.close()


@pytest.fixture(scope="module")
def served_full(ctx):
    wl = S.serve_workload()
    gs = GpuServeScene(ctx, wl)
    yield wl, gs
    gs.close()


def test_served_fullsize_mix_equals_alone(ctx, served_full):
    wl, gs = served_full
    views = wl.requests(64, seed=3)
    big = max(range(len(wl.meshes)), key=lambda m: wl.meshes[m][1].shape[0])
    views.append(S.ViewSpec(mesh=big, width=1920, height=1080, camera_pos=S.orbit_position(2.6, 40.0, 20.0)))
    with Server(ctx, lanes=3, max_batch_views=16, with_ids=True) as srv:
        srv.pause(True)     # the whole stream queues, then is admitted in shared batches
        tickets = [srv.submit(gs.meshes[v.mesh], v.width, v.height, gs.params(v)) for v in views]
        srv.pause(False)
        for t in tickets:
            srv.wait(t, 120.0)
        results = [srv.read(t) for t in tickets]
        st = srv.stats()
        for t in tickets:
            srv.release(t)
    assert st["completed"] == len(views) and st["failed"] == 0, st
    assert st["max_views_in_batch"] > 1, st
    assert st["overflow_retries"] == 0, st
    for i, v in enumerate(views):
        alone = gs.draw_plain(v, with_ids=True)
        for a, b, name in zip(results[i], alone, ("colour", "depth", "ids")):
            assert (a.view(np.uint8) == b.view(np.uint8)).all(), f"request {i}: {name} differs from alone"
    # the checker on a sample: the largest mesh at 1080p and one of each resolution
    sample = [len(views) - 1]
    for w, h, _ in S.SERVE_RESOLUTIONS:
        sample.append(next(i for i, v in enumerate(views) if (v.width, v.height) == (w, h)))
    for i in sample:
        rep = H.compare(results[i], H.render_oracle(wl.scene(views[i])))
        H.assert_parity(rep, f"served request {i} (mesh {views[i].mesh}, {views[i].width}x{views[i].height})")
        assert rep["covered"] > 0
