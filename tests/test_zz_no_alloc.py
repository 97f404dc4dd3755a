"""axr_draw_noalloc: drawMesh with the frame's working memory in one fixed
device array instead of allocated frame arenas.  Its frames must equal
axr_draw's bit for bit (the same kernels on the same inputs), on two calls
in a row with the second smaller than the first, synchronous and
asynchronous, with and without a recorded clear; a frame larger than the
array is refused before any work is queued."""
import ctypes as C

import numpy as np
import pytest

import harness as H
from axiomr_amd import _native as N
from axiomr_amd import render as R
from axiomr_amd import scenes as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    """A context of this module's own, so that these tests run last (a
    session-wide parametrized context would reorder them among the others);
    axr_draw is compared in both raster modes."""
    import axiomr_amd as axr
    c = axr.Context(0)
    yield c
    c.close()


@pytest.fixture(params=[0, 1], ids=["binned", "visibility"])
def mode(request, ctx):
    ctx.set_config(raster_mode=request.param)
    return request.param


def _draw(ctx, fn, fb, gs, p, clear):
    if clear:
        fb.clear(R.Color(9, 20, 30, 255))
    code = fn(ctx.handle, fb.handle, gs.mesh.handle, C.byref(p))
    ctx.check(code)
    ctx.sync()
    return fb.read()


def _same(a, b, tag):
    for x, y, name in zip(a, b, ("colour", "depth", "ids")):
        if x is None and y is None:
            continue
        assert (x.view(np.uint8) == y.view(np.uint8)).all(), f"{tag}: {name} differs from axr_draw"


@pytest.mark.parametrize("flags", [0, N.AXR_DRAW_ASYNC], ids=["sync", "async"])
@pytest.mark.parametrize("with_ids", [True, False])
def test_noalloc_equals_draw_twice_shrinking(ctx, mode, flags, with_ids):
    L = ctx._lib
    scenes = [S.icosphere_scene(n=96, width=640, height=480, tex=256),
              S.icosphere_scene(n=24, width=200, height=150, tex=64, yaw=70.0)]
    for k, sc in enumerate(scenes):          # the second call is the smaller one
        gs = H.GpuScene(ctx, sc)
        fa = R.Framebuffer(ctx, sc.width, sc.height, with_ids=with_ids)
        fb = R.Framebuffer(ctx, sc.width, sc.height, with_ids=with_ids)
        try:
            for clear in (True, False):      # a recorded clear absorbed, then a draw on top
                p = H.params_of(sc, flags=flags)
                ref = _draw(ctx, L.axr_draw, fa, gs, p, clear)
                got = _draw(ctx, L.axr_draw_noalloc, fb, gs, p, clear)
                _same(got, ref, f"scene {k} clear={clear}")
                assert int(np.isfinite(got[1]).sum()) > 1000
            # and the checker agrees with the first frame drawn
            if with_ids:
                fb.clear(H.BLACK)
                got = _draw(ctx, L.axr_draw_noalloc, fb, gs, H.params_of(sc, flags=flags), False)
                H.assert_parity(H.compare(got, H.render_oracle(sc)), f"noalloc scene {k}")
        finally:
            fa.close()
            fb.close()
            gs.close()


def test_noalloc_refuses_a_frame_larger_than_its_scratch(ctx, mode):
    sc = S.icosphere_scene(n=8, width=64, height=64, tex=16)
    gs = H.GpuScene(ctx, sc)
    big = R.Framebuffer(ctx, 4096, 4096)     # 8 bytes of keys a pixel: 128 MiB > 64 MiB
    try:
        code = ctx._lib.axr_draw_noalloc(ctx.handle, big.handle, gs.mesh.handle, C.byref(H.params_of(sc)))
        assert code == N.AXR_ERR_INVALID
        assert b"scratch" in ctx._lib.axr_last_error(ctx.handle)
        ctx.sync()
    finally:
        big.close()
        gs.close()
