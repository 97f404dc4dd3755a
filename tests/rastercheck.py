"""Per-pixel brute force of the reference raster's coverage and depth rule
(reference src/tiled_pipeline.cpp:436-594 and 1125-1181), in numpy float32,
evaluated for every (pixel, triangle) pair with no tiles, bins, threads or
merge - the structure the CPU checker and the GPU path are built from.  Each
numpy operation rounds once to float32, so the evaluation order written
here is the one computed."""
from __future__ import annotations

import numpy as np

f32 = np.float32
EMPTY = 0xFFFFFFFF


def screen_setup(clip, W, H):
    """clip: (T, 3, 4) float32 -> screen x, y (T, 3) and ndc z (T, 3)."""
    w = clip[..., 3].copy()
    tiny = f32(1e-6)
    small = np.abs(w) < tiny
    w[small] = np.where(w[small] < 0, -tiny, tiny)
    inv = f32(1) / w
    sx = (((clip[..., 0] * inv) + f32(1)) * f32(0.5)) * f32(W)
    sy = (((clip[..., 1] * inv) + f32(1)) * f32(0.5)) * f32(H)
    z = clip[..., 2] * inv
    return sx.astype(f32), sy.astype(f32), z.astype(f32)


def backface(sx, sy):
    dx1, dy1 = sx[:, 1] - sx[:, 0], sy[:, 1] - sy[:, 0]
    dx2, dy2 = sx[:, 2] - sx[:, 0], sy[:, 2] - sy[:, 0]
    return (dx1 * dy2 - dx2 * dy1) < 0


def first_min(a, axis):
    """min keeping the first of equal values, as std::min({a,b,c}) does."""
    r = a[:, 0].copy()
    for k in (1, 2):
        r = np.where(a[:, k] < r, a[:, k], r)
    return r


def first_max(a, axis):
    r = a[:, 0].copy()
    for k in (1, 2):
        r = np.where(r < a[:, k], a[:, k], r)
    return r


def to_int(f):
    out = np.full(f.shape, np.iinfo(np.int32).min, np.int64)
    ok = (f >= -2147483648.0) & (f < 2147483648.0)
    out[ok] = np.trunc(f[ok]).astype(np.int64)
    return out


def resolve(sx, sy, z, keys, W, H):
    """Winner per pixel: argmin over covered triangles of (z, key), with -0
    equal to +0 and NaN/+inf never winning.  Returns depth (H, W) float32
    (+inf where empty) and ids (H, W) uint32."""
    T = sx.shape[0]
    depth = np.full((H, W), np.inf, f32)
    ids = np.full((H, W), EMPTY, np.uint32)
    py, px = np.mgrid[0:H, 0:W]
    px = px.astype(np.int64)
    py = py.astype(np.int64)
    minx, maxx = first_min(sx, 1), first_max(sx, 1)
    miny, maxy = first_min(sy, 1), first_max(sy, 1)
    fx0, fy0 = to_int(np.floor(minx)), to_int(np.floor(miny))
    cx1, cy1 = to_int(np.ceil(maxx)), to_int(np.ceil(maxy))
    for t in range(T):
        x0, y0, x1, y1, x2, y2 = sx[t, 0], sy[t, 0], sx[t, 1], sy[t, 1], sx[t, 2], sy[t, 2]
        c0 = x1 * y2 - x2 * y1
        c1 = x2 * y0 - x0 * y2
        c2 = x0 * y1 - x1 * y0
        area = (c0 + c1) + c2
        if area >= 0 and float(area) < 1.0e-12:
            continue
        a0, b0 = y1 - y2, x2 - x1
        a1, b1 = y2 - y0, x0 - x2
        a2, b2 = y0 - y1, x1 - x0
        if area < 0:
            a0, b0, c0, a1, b1, c1, a2, b2, c2 = -a0, -b0, -c0, -a1, -b1, -c1, -a2, -b2, -c2
            area = -area
        inv = f32(1) / area
        inside = (px >= fx0[t]) & (px < cx1[t]) & (py >= fy0[t]) & (py < cy1[t])
        if not inside.any():
            continue
        sxa = np.maximum((px // 16) * 16, fx0[t])          # 16-pixel tile anchor
        sxc = sxa.astype(f32) + f32(0.5)
        pyc = py.astype(f32) + f32(0.5)
        d = px - sxa
        lane = (d % 8).astype(f32)
        step = d >= 8
        es = []
        for a, b, c in ((a0, b0, c0), (a1, b1, c1), (a2, b2, c2)):
            row = (a * sxc + b * pyc) + c
            row = np.where(step, row + a * f32(8), row)
            es.append(row + a * lane)
        cov = inside & (es[0] >= 0) & (es[1] >= 0) & (es[2] >= 0)
        al, be, ga = es[0] * inv, es[1] * inv, es[2] * inv
        zz = (z[t, 0] * al + z[t, 1] * be) + z[t, 2] * ga
        # strict-less against the current winner; equal depth keeps the
        # earlier key (triangles arrive here in key order)
        better = cov & (zz < depth)
        depth = np.where(better, zz, depth)
        ids = np.where(better, np.uint32(keys[t]), ids)
    return depth.astype(f32), ids


def merge_draws(draws, W, H, clear_bgra, clear_depth):
    """Framebuffer after a clear and a list of draws, each draw given as
    (sx, sy, z, keys, bgra per triangle): the draw's own winners
    (`resolve`, from +inf as the reference's tile buffers start), then the
    reference's merge (tiled_pipeline.cpp:1125-1181): a pixel takes the
    draw's depth and colour only if it is strictly less than what the
    framebuffer holds, so an earlier draw or the clear keeps every tie.
    Returns colour (H, W) uint32 BGRA-packed, depth (H, W) float32, and the
    (draw index, key) of each written pixel (-1, EMPTY where the clear
    stayed)."""
    color = np.full((H, W), np.uint32(clear_bgra), np.uint32)
    depth = np.full((H, W), f32(clear_depth), f32)
    src = np.full((H, W), -1, np.int64)
    ids = np.full((H, W), EMPTY, np.uint32)
    for d, (sx, sy, z, keys, bgra) in enumerate(draws):
        dz, did = resolve(sx, sy, z, keys, W, H)
        win = (did != EMPTY) & (dz < depth)
        slot = np.searchsorted(keys, did[win])
        color[win] = np.asarray(bgra, np.uint32)[slot]
        depth[win] = dz[win]
        src[win] = d
        ids[win] = did[win]
    return color, depth, src, ids
