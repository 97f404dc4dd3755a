"""The batched form: many views of several resident meshes drawn by one
graph-captured launch must give, view by view, exactly what the plain
per-frame path gives alone (and what the CPU checker gives)."""
import numpy as np
import pytest

import harness as H
from axiomr_amd import scenes as S

pytestmark = pytest.mark.gpu


def small_workload(seed=5):
    return S.batch_workload(nviews=24, nmeshes=6, fmin=500, fmax=40_000, wmin=40, wmax=400, hmin=32, hmax=300,
                            tex=128, seed=seed)


@pytest.mark.parametrize("tile", [16, 8, 32])
def test_batch_equals_solo(ctx, tile):
    wl = small_workload()
    gb = H.GpuBatch(ctx, wl, tile=tile)
    try:
        gb.batch.launch()
        ctx.sync()
        batched = [fb.read() for fb in gb.fbs]
        gb.batch.launch()        # a replay gives the same frame
        ctx.sync()
        again = [fb.read() for fb in gb.fbs]
        for i in range(len(wl.views)):
            gb.draw_solo(i)
            solo = gb.fbs[i].read()
            for a, b, c in zip(batched[i], solo, again[i]):
                assert (a.view(np.uint8) == b.view(np.uint8)).all(), f"view {i} differs from its solo render"
                assert (a.view(np.uint8) == c.view(np.uint8)).all(), f"view {i} differs between replays"
        covered = sum(int(np.isfinite(b[1]).sum()) for b in batched)
        assert covered > 50_000
    finally:
        gb.close()


def test_batch_views_match_checker(ctx):
    wl = small_workload(seed=9)
    gb = H.GpuBatch(ctx, wl)
    try:
        gb.batch.launch()
        ctx.sync()
        for i in range(0, len(wl.views), 4):
            rep = H.compare(gb.fbs[i].read(), H.render_oracle(wl.scene(i)))
            H.assert_parity(rep, f"batch view {i}")
    finally:
        gb.close()


def test_batch_rejects_a_framebuffer_in_two_views(ctx):
    """Every view of a batch resolves at the same time, so two views drawing
    into one framebuffer would race on its pixels: creation refuses it with
    AXR_ERR_INVALID (and accepts the same views on two framebuffers)."""
    from axiomr_amd import render as R
    from axiomr_amd._native import AXR_ERR_INVALID, AxrError
    sc = S.icosphere_scene(n=8, width=64, height=48, tex=16)
    gs = H.GpuScene(ctx, sc)
    fbs = [R.Framebuffer(ctx, sc.width, sc.height) for _ in range(2)]
    p = H.params_of(sc)
    try:
        with pytest.raises(AxrError) as ei:
            R.Batch(ctx, [fbs[0], fbs[0]], [gs.mesh, gs.mesh], [p, p])
        assert ei.value.code == AXR_ERR_INVALID
        ok = R.Batch(ctx, fbs, [gs.mesh, gs.mesh], [p, p])
        ok.close()
    finally:
        for fb in fbs:
            fb.close()
        gs.close()
