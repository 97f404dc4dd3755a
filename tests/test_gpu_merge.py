"""Cross-draw merge on the GPU: two draws of the same triangles with
different flat colours into one framebuffer, after a clear whose depth may
equal a winner's exact depth.  The reference merges each draw into the
framebuffer with a strict-less depth test (tiled_pipeline.cpp:1125-1181,
line 1150), so every tie keeps the earlier draw's colour, and a pixel whose
winner sits exactly at the clear depth keeps the clear.  Checked against
the CPU checker (itself pinned by the brute force in rastercheck.py) for
synchronous draws, asynchronous draws (the two-pass resolve that absorbs
the clear) and a mix, in both raster modes."""
import numpy as np
import pytest

from axiomr_amd import render as R
from axiomr_amd.resident import GpuScene
from test_oracle_raster import _oracle_draws, identity_params, ndc_scene, random_tris

pytestmark = pytest.mark.gpu

W, HH = 96, 64


def _scenes():
    rng = np.random.default_rng(33)
    tris = random_tris(rng, 60, 0.3)
    ca = rng.integers(0, 2**32, 60, dtype=np.uint64).astype(np.uint32)
    cb = rng.integers(0, 2**32, 60, dtype=np.uint64).astype(np.uint32)
    return ndc_scene(tris, W, HH, colors=ca), ndc_scene(tris, W, HH, colors=cb)


@pytest.mark.parametrize("flags", [(0, 0), (1, 1), (1, 0), (0, 1)], ids=["sync", "async", "async-sync", "sync-async"])
@pytest.mark.parametrize("clear_at", ["inf", "winner"])
def test_first_draw_wins_ties(ctx, flags, clear_at):
    sa, sb = _scenes()
    clear_depth = np.inf
    if clear_at == "winner":
        _, d0, _ = _oracle_draws([sa], W, HH, np.inf)
        fin = np.sort(d0[np.isfinite(d0)])
        clear_depth = float(fin[len(fin) // 2])
    ref = _oracle_draws([sa, sb], W, HH, clear_depth)
    ga, gb = GpuScene(ctx, sa), GpuScene(ctx, sb)
    fb = R.Framebuffer(ctx, W, HH, with_ids=True)
    try:
        fb.clear(R.Color(0, 0, 0, 255), clear_depth)
        for g, f in zip((ga, gb), flags):
            p = identity_params()
            p.flags = f
            ctx.check(ctx._lib.axr_draw(ctx.handle, fb.handle, g.mesh.handle, R.C.byref(p)))
        c, d, ids = fb.read()
    finally:
        fb.close()
        ga.close()
        gb.close()
    assert (d.view(np.uint32) == ref[1].view(np.uint32)).all(), "depth bits differ"
    assert (ids == ref[2]).all(), "winners differ"
    assert (c == ref[0]).all(), "colours differ: a later draw won a tie"
    # the second draw changed no pixel: what the first draw alone gives
    alone = _oracle_draws([sa], W, HH, clear_depth)
    assert (c == alone[0]).all()
    if clear_at == "winner":
        _, d0, _ = _oracle_draws([sa], W, HH, np.inf)
        at = d0 == np.float32(clear_depth)
        assert at.any() and (c[at] == [0, 0, 0, 255]).all()
