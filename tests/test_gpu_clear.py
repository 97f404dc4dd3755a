"""A framebuffer clear is recorded by axr_fb_clear and written by the next
draw's last pass (visibility mode) or before any other use of the planes
(read-back, binned draw, batch).  Whatever the order of calls, the pixels
are those of the clear followed by the draws, as the CPU checker gives them."""
import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import harness as H
import oracle as O
from axiomr_amd import _native as N
from axiomr_amd import render as R
from axiomr_amd import scenes as S

pytestmark = pytest.mark.gpu

TINT = R.Color(10, 200, 30, 77)


def small_scene():
    return S.icosphere_scene(n=16, width=200, height=150, tex=64)


def oracle_sequence(scene, steps):
    """steps: ("clear", color, depth) or ("draw", params) applied in order."""
    os_ = O.OracleScene(scene)
    rr = O.OracleRenderer(0)
    fb = rr.framebuffer(scene.width, scene.height)
    try:
        O.fb_clear(fb, 0, float("inf"))
        for st in steps:
            if st[0] == "clear":
                O.fb_clear(fb, st[1].packed_bgra(), st[2])
            else:
                rr.draw(fb, os_, st[1])
        return O.fb_read(fb, scene.width, scene.height)
    finally:
        O.fb_destroy(fb)
        rr.close()
        os_.close()


def draw(ctx, fb, gs, p):
    ctx.check(ctx._lib.axr_draw(ctx.handle, fb.handle, gs.mesh.handle, C.byref(p)))


def median_depth(scene):
    d = oracle_sequence(scene, [("draw", H.params_of(scene))])[1]
    return float(np.median(d[np.isfinite(d)]))


def test_clear_colour_and_depth_then_draw(ctx):
    sc = small_scene()
    d = median_depth(sc)                      # half the sphere's pixels lie behind the clear
    gs = H.GpuScene(ctx, sc)
    fb = R.Framebuffer(ctx, sc.width, sc.height, with_ids=True)
    try:
        p = H.params_of(sc)
        fb.clear(TINT, d)
        draw(ctx, fb, gs, p)
        got = fb.read()
    finally:
        fb.close()
        gs.close()
    ref = oracle_sequence(sc, [("clear", TINT, d), ("draw", p)])
    H.assert_parity(H.compare(got, ref), "clear at the median depth")
    assert (got[1] == np.float32(d)).sum() > 1000 and (got[1] < d).sum() > 1000


def test_clear_alone_reads_back(ctx):
    sc = small_scene()
    gs = H.GpuScene(ctx, sc)
    fb = R.Framebuffer(ctx, sc.width, sc.height, with_ids=True)
    try:
        fb.clear(TINT, 0.25)
        bgra, depth, ids = fb.read()
        assert (bgra.view(np.uint32) == TINT.packed_bgra()).all() and (depth == np.float32(0.25)).all()
        assert (ids == 0xFFFFFFFF).all()
        draw(ctx, fb, gs, H.params_of(sc))
        fb.clear(H.BLACK)                     # a clear after a draw wipes it
        bgra, depth, ids = fb.read()
        assert (bgra.view(np.uint32) == H.BLACK.packed_bgra()).all() and np.isinf(depth).all()
        assert (ids == 0xFFFFFFFF).all()
    finally:
        fb.close()
        gs.close()


def test_last_clear_wins_and_second_draw_sees_the_first(ctx):
    sc = small_scene()
    gs = H.GpuScene(ctx, sc)
    fb = R.Framebuffer(ctx, sc.width, sc.height, with_ids=True)
    p1 = H.params_of(sc)
    a = np.radians(35.0)
    rot = np.array([[np.cos(a), 0, np.sin(a), 0], [0, 1, 0, 0], [-np.sin(a), 0, np.cos(a), 0], [0, 0, 0, 1]])
    p2 = H.params_of(sc, model=rot.T.astype(np.float32).reshape(16))   # column-major
    try:
        fb.clear(R.Color(1, 2, 3, 4), 0.1)
        fb.clear(TINT)
        draw(ctx, fb, gs, p1)
        draw(ctx, fb, gs, p2)
        got = fb.read()
    finally:
        fb.close()
        gs.close()
    ref = oracle_sequence(sc, [("clear", TINT, float("inf")), ("draw", p1), ("draw", p2)])
    H.assert_parity(H.compare(got, ref), "two draws after one clear")


def test_async_draw_after_clear(ctx):
    sc = small_scene()
    gs = H.GpuScene(ctx, sc)
    fb = R.Framebuffer(ctx, sc.width, sc.height, with_ids=True)
    try:
        p = H.params_of(sc, flags=N.AXR_DRAW_ASYNC)
        fb.clear(TINT)
        draw(ctx, fb, gs, p)
        ctx.sync()
        got = fb.read()
    finally:
        fb.close()
        gs.close()
    p.flags = 0
    H.assert_parity(H.compare(got, oracle_sequence(sc, [("clear", TINT, float("inf")), ("draw", p)])), "async")


def test_batch_without_its_own_clear_writes_pending_clears(ctx):
    sc = small_scene()
    gs = H.GpuScene(ctx, sc)
    fbs = [R.Framebuffer(ctx, sc.width, sc.height, with_ids=True) for _ in range(3)]
    params = [H.params_of(sc) for _ in range(3)]
    batch = R.Batch(ctx, fbs, [gs.mesh] * 3, params)
    try:
        batch.set_clear(None)
        for i, fb in enumerate(fbs):
            fb.clear(R.Color(10 * i, 20, 30, 255), 1.0 if i == 1 else float("inf"))
        batch.launch()
        ctx.sync()
        for i, fb in enumerate(fbs):
            ref = oracle_sequence(sc, [("clear", R.Color(10 * i, 20, 30, 255), 1.0 if i == 1 else float("inf")),
                                       ("draw", params[i])])
            H.assert_parity(H.compare(fb.read(), ref), f"batch view {i}")
        # the batch's own clear replaces a pending one
        batch.set_clear(TINT)
        fbs[0].clear(R.Color(1, 1, 1, 1), 0.0)
        batch.launch()
        ctx.sync()
        ref = oracle_sequence(sc, [("clear", TINT, float("inf")), ("draw", params[0])])
        H.assert_parity(H.compare(fbs[0].read(), ref), "batch clear over a pending clear")
    finally:
        batch.close()
        for fb in fbs:
            fb.close()
        gs.close()


def quad_stack(n=20, width=1920, height=1080):
    """n screen-filling quads (2n triangles, larger than the view, so they are
    clipped) at depths 1.0 .. 1.0 + 0.1 (n - 1), flat per-face colours."""
    v = np.zeros((4 * n, 14), np.float32)
    f = np.zeros((2 * n, 3), np.uint32)
    for k in range(n):
        z = -(1.0 + 0.1 * k)
        s = 4.0 * -z
        v[4 * k:4 * k + 4, :3] = [(-s, -s, z), (s, -s, z), (s, s, z), (-s, s, z)]
        v[4 * k:4 * k + 4, 5:8] = (0.0, 0.0, 1.0)
        f[2 * k] = (4 * k, 4 * k + 1, 4 * k + 2)
        f[2 * k + 1] = (4 * k, 4 * k + 2, 4 * k + 3)
    rng = np.random.default_rng(4)
    colors = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    return S.SceneData(name="quad_stack", vertices=v, faces=f, width=width, height=height,
                       camera_pos=(0.0, 0.0, 0.0), camera_target=(0.0, 0.0, -1.0), fov=60.0, shader={"kind": 5},
                       groups=[(0, 0, f.shape[0])], materials=[S.MaterialData()], face_colors=colors)


@pytest.mark.parametrize("mode,with_ids", [(N.RASTER_BINNED, True), (N.RASTER_VISIBILITY, True),
                                           (N.RASTER_VISIBILITY, False)])
def test_overflowing_frame_after_clear(mode, with_ids):
    """A fresh context's arena is too small for 40 clipped screen-filling
    triangles at 1080p: a synchronous draw grows it and redraws; an
    asynchronous one reports the overflow at sync and leaves the cleared
    framebuffer, as if its clear had run on its own (without an id plane the
    asynchronous draw shades and writes in two passes)."""
    sc = quad_stack()
    p = H.params_of(sc)
    ref = oracle_sequence(sc, [("clear", TINT, float("inf")), ("draw", p)])
    for asynchronous in (False, True):
        ctx = R.Context(0)
        try:
            ctx.set_config(tile_size=16, raster_threads=256, small_footprint=128, raster_mode=mode)
            gs = H.GpuScene(ctx, sc)
            fb = R.Framebuffer(ctx, sc.width, sc.height, with_ids=with_ids)
            try:
                fb.clear(TINT)
                if asynchronous:
                    q = H.params_of(sc, flags=N.AXR_DRAW_ASYNC)
                    ctx.check(ctx._lib.axr_draw(ctx.handle, fb.handle, gs.mesh.handle, C.byref(q)))
                    assert ctx._lib.axr_sync(ctx.handle) == -4          # AXR_ERR_OVERFLOW
                    bgra, depth, ids = fb.read()
                    assert (bgra.view(np.uint32) == TINT.packed_bgra()).all() and np.isinf(depth).all()
                    assert ids is None if not with_ids else (ids == 0xFFFFFFFF).all()
                    fb.clear(TINT)
                draw(ctx, fb, gs, p)                                    # grows (again) if needed, redraws
                got = fb.read()
                H.assert_parity(H.compare(got, ref if with_ids else (ref[0], ref[1], None)),
                                f"quad stack mode {mode} async {asynchronous}")
            finally:
                fb.close()
                gs.close()
        finally:
            ctx.close()


def rotated(sc, deg, flags=0):
    a = np.radians(deg)
    rot = np.array([[np.cos(a), 0, np.sin(a), 0], [0, 1, 0, 0], [-np.sin(a), 0, np.cos(a), 0], [0, 0, 0, 1]])
    return H.params_of(sc, model=rot.T.astype(np.float32).reshape(16), flags=flags)


def test_async_draws_keep_call_order(ctx):
    """Asynchronous draws take the frame arenas in turn and overlap; draws
    into one framebuffer must still land in call order: six rotated
    copies drawn without clears in between (each depth-tested against the
    ones before), interleaved with draws into a second framebuffer, a
    synchronous draw and a clear."""
    sc = small_scene()
    gs = H.GpuScene(ctx, sc)
    fa = R.Framebuffer(ctx, sc.width, sc.height, with_ids=True)
    fb = R.Framebuffer(ctx, sc.width, sc.height, with_ids=True)
    steps_a, steps_b = [("clear", TINT, float("inf"))], [("clear", H.BLACK, float("inf"))]
    try:
        fa.clear(TINT)
        fb.clear(H.BLACK)
        for k in range(6):
            p = rotated(sc, 17.0 * k, N.AXR_DRAW_ASYNC)
            draw(ctx, fa, gs, p)
            steps_a.append(("draw", rotated(sc, 17.0 * k)))
            if k % 2 == 0:
                q = rotated(sc, -23.0 * k, N.AXR_DRAW_ASYNC if k != 2 else 0)
                if k == 4:
                    fb.clear(R.Color(5, 6, 7, 8), 2.0)
                    steps_b.append(("clear", R.Color(5, 6, 7, 8), 2.0))
                draw(ctx, fb, gs, q)
                steps_b.append(("draw", rotated(sc, -23.0 * k)))
        ctx.sync()
        got_a, got_b = fa.read(), fb.read()
    finally:
        fa.close()
        fb.close()
        gs.close()
    H.assert_parity(H.compare(got_a, oracle_sequence(sc, steps_a)), "six async draws into one framebuffer")
    H.assert_parity(H.compare(got_b, oracle_sequence(sc, steps_b)), "interleaved framebuffer")


@pytest.mark.parametrize("width,height,tile", [(200, 150, 8), (333, 197, 8), (333, 197, 16), (256, 130, 32),
                                               (250, 130, 32)])
def test_async_frames_without_ids_shade_ahead(width, height, tile):
    """Asynchronous clear-then-draw frames into framebuffers without an id
    plane shade before the earlier frames have written and write after
    them: frames round-robin over three framebuffers, a clear at the median
    depth (winners behind it are dropped), a second draw without a clear
    between (the one-pass resolve, depth-tested against what the two-pass
    one wrote), and other shader kinds.  Widths that are and are not a
    multiple of 4 (16-byte and 4-byte framebuffer writes) and partial bins."""
    ctx = R.Context(0)
    try:
        ctx.set_config(tile_size=tile, raster_mode=N.RASTER_VISIBILITY)
        _frames_without_ids(ctx, S.icosphere_scene(n=16, width=width, height=height, tex=64))
    finally:
        ctx.close()


@pytest.mark.parametrize("width,height,tile", [(200, 150, 16), (333, 197, 8), (250, 130, 32)])
def test_async_frames_with_ids_two_pass(width, height, tile):
    """The same frames into framebuffers with an id plane: the two-pass
    resolve keeps each winner's key in its slot and its colour in the
    arena's colour slots, and the finishing pass writes all three planes
    (16-byte writes when the width is a multiple of 4, else 4-byte)."""
    ctx = R.Context(0)
    try:
        ctx.set_config(tile_size=tile, raster_mode=N.RASTER_VISIBILITY)
        _frames_without_ids(ctx, S.icosphere_scene(n=16, width=width, height=height, tex=64), with_ids=True)
    finally:
        ctx.close()


def _frames_without_ids(ctx, sc, with_ids=False):
    d = median_depth(sc)
    light = (0.2, -1.0, -0.3)
    variants = {4: sc,
                0: S.with_shader(sc, {"kind": 0, "lightDirection": light}),
                1: S.with_shader(sc, {"kind": 1, "lightDirection": light, "lightColor": (0.6, 0.6, 0.6)}),
                3: S.with_shader(sc, {"kind": 3, "lightDirection": light})}
    gs = H.GpuScene(ctx, sc)
    fbs = [R.Framebuffer(ctx, sc.width, sc.height, with_ids=with_ids) for _ in range(3)]
    steps = [[] for _ in fbs]
    try:
        for i, kind in enumerate([4, 0, 1, 3, 4, 4, 1, 4, 0]):
            k, v = i % 3, variants[kind]
            col, dep = R.Color(20 * i, 7, 99, 255), (d if i == 4 else float("inf"))
            fbs[k].clear(col, dep)
            steps[k] = [("clear", col, dep)]
            draw(ctx, fbs[k], gs, rotated(v, 11.0 * i, N.AXR_DRAW_ASYNC))
            steps[k].append(("draw", rotated(v, 11.0 * i)))
            if i == 6:
                draw(ctx, fbs[k], gs, rotated(v, -40.0, N.AXR_DRAW_ASYNC))
                steps[k].append(("draw", rotated(v, -40.0)))
        ctx.sync()
        got = [fb.read() for fb in fbs]
    finally:
        for fb in fbs:
            fb.close()
        gs.close()
    for k in range(3):
        ref = oracle_sequence(sc, steps[k])
        H.assert_parity(H.compare(got[k], ref if with_ids else (ref[0], ref[1], None)), f"framebuffer {k}")
        if with_ids:
            assert (got[k][2] != H.EMPTY_ID).sum() > 100


@pytest.mark.parametrize("tile", [8, 16])
def test_async_soup_with_clipping_and_ties_two_pass(tile):
    """Shading ahead of the framebuffer pass on a scene with clipped faces
    (their pieces read back through the clip records) and exactly equal
    depths (submission order decides): queued clear-then-draw frames into a
    framebuffer without an id plane, two of them into the same framebuffer
    back to back.  A fresh context's arenas are too small for this scene's
    large triangles: the first sync reports the overflow, every arena is
    grown to it, and the same frames drawn again succeed."""
    sc = S.soup_scene(ntri=60_000, width=480, height=270, seed=9)
    ref = H.render_oracle(sc)
    ctx = R.Context(0)
    try:
        ctx.set_config(tile_size=tile, raster_mode=N.RASTER_VISIBILITY)
        gs = H.GpuScene(ctx, sc)
        fbs = [R.Framebuffer(ctx, sc.width, sc.height) for _ in range(2)]
        try:
            p = H.params_of(sc, flags=N.AXR_DRAW_ASYNC)
            overflows = 0
            while True:
                for fb in (fbs[0], fbs[1], fbs[0]):
                    fb.clear(H.BLACK)
                    draw(ctx, fb, gs, p)
                if ctx._lib.axr_sync(ctx.handle) != N.AXR_ERR_OVERFLOW:
                    break
                overflows += 1
                assert overflows < 2, "arenas not grown to the reported needs"
            ctx.check(ctx._lib.axr_sync(ctx.handle))
            got = [fb.read() for fb in fbs]
        finally:
            for fb in fbs:
                fb.close()
            gs.close()
    finally:
        ctx.close()
    for k, g in enumerate(got):
        rep = H.compare(g, (ref[0], ref[1], None))
        H.assert_parity(rep, f"soup tile {tile} framebuffer {k}")
        assert rep["covered"] > 50_000

_CHILD = """
import sys
sys.path[:0] = ["tests", "."]
import test_gpu_clear as T
import axiomr_amd as axr
from axiomr_amd import _native as N
from axiomr_amd import scenes as S
ctx = axr.Context(0)
ctx.set_config(tile_size=16, raster_mode=N.RASTER_VISIBILITY)
T._frames_without_ids(ctx, S.icosphere_scene(n=16, width=200, height=150, tex=64))
T.test_async_draws_keep_call_order(ctx)
ctx.close()
print("ok")
"""


@pytest.mark.parametrize("env", [{"AXR_STREAM_WAIT_VALUE": "0"}, {"AXR_OVERLAP": "0"}])
def test_ordering_with_handoff_fallbacks(env):
    """The ordering checks above in a process whose overlapped draws hand
    the framebuffer over with events instead of stream memory waits, and in
    one whose draws do not overlap (what a counter-collecting profiler
    gets).  Both switches are read once per process, hence the child."""
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, "-c", _CHILD], cwd=root, env={**os.environ, **env}, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-3000:] + r.stderr[-3000:]
