"""The CPU checker's triangle setup, cull and clipper against closed forms
and invariants (reference: src/tiled_pipeline.cpp:15-119, src/pipeline.cpp
:176-370).  A plausible slip - a wrong sign in a plane, a swapped rotation,
a dropped perspective term in the uv interpolation - fails one of these."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

F = C.POINTER(C.c_float)


def setup(clip12, W, H):
    out = np.zeros(13, np.float32)
    c = np.ascontiguousarray(clip12, np.float32)
    back = O.lib().orc_setup_triangle(c.ctypes.data_as(F), W, H, out.ctypes.data_as(F))
    return out, bool(back)


def clip(tri18):
    t = np.ascontiguousarray(tri18, np.float32).reshape(3, 18)
    out = np.zeros((24, 18), np.float32)
    n = O.lib().orc_clip_triangle(t.ctypes.data_as(F), out.ctypes.data_as(F))
    return out[:n]


def clip_plane(plane, tri18):
    t = np.ascontiguousarray(tri18, np.float32).reshape(3, 18)
    out = np.zeros((4, 18), np.float32)
    r = O.lib().orc_clip_single_plane(plane, t.ctypes.data_as(F), out.ctypes.data_as(F))
    return r, out


def test_setup_matches_closed_form():
    rng = np.random.default_rng(1)
    W, H = 1920, 1080
    for _ in range(500):
        c = rng.uniform(-1, 1, (3, 4)).astype(np.float32)
        c[:, 3] = rng.uniform(0.2, 5.0, 3)
        out, back = setup(c.reshape(-1), W, H)
        c64 = c.astype(np.float64)
        sx = (c64[:, 0] / c64[:, 3] + 1) * 0.5 * W
        sy = (c64[:, 1] / c64[:, 3] + 1) * 0.5 * H
        z = c64[:, 2] / c64[:, 3]
        got = out[:6].reshape(3, 2).astype(np.float64)
        assert np.allclose(got[:, 0], sx, rtol=3e-6, atol=3e-4)
        assert np.allclose(got[:, 1], sy, rtol=3e-6, atol=3e-4)
        assert np.allclose(out[6:9], z, rtol=3e-6, atol=1e-6)
        assert out[9] == got[:, 0].min() and out[11] == got[:, 0].max()
        assert out[10] == got[:, 1].min() and out[12] == got[:, 1].max()
        area = (sx[1] - sx[0]) * (sy[2] - sy[0]) - (sx[2] - sx[0]) * (sy[1] - sy[0])
        if abs(area) > 1e-2:
            assert back == (area < 0), "counter-clockwise (y up) is front-facing"


def test_setup_clamps_tiny_w():
    # |w| < 1e-6 is replaced by +-1e-6 before the divide (tiled_pipeline.cpp:15-23)
    c = np.array([[1e-7, 0, 0, 1e-7], [0.5, 0.5, 0, 1], [0, 0.5, 0, 1]], np.float32)
    out, _ = setup(c.reshape(-1), 100, 100)
    assert out[0] == np.float32(((np.float32(1e-7) * np.float32(1e6)) + 1) * 0.5 * 100)
    c[0, 3] = -1e-8
    c[0, 0] = -1e-7
    out, _ = setup(c.reshape(-1), 100, 100)
    assert out[0] == np.float32(((np.float32(-1e-7) * (np.float32(1) / np.float32(-1e-6))) + 1) * 0.5 * 100)


def rand_vertex(rng, clip4):
    v = np.zeros(18, np.float32)
    v[:3] = rng.normal(size=3)
    v[3:5] = rng.uniform(0, 1, 2)
    for k in (5, 8, 11):
        d = rng.normal(size=3)
        v[k:k + 3] = d / np.linalg.norm(d)
    v[14:18] = clip4
    return v


def dist(c, plane):
    x, y, z, w = c
    return [x + w, w - x, y + w, w - y, z + w, w - z][plane]


def test_inside_triangle_passes_through_bit_exact():
    rng = np.random.default_rng(2)
    for _ in range(100):
        t = np.stack([rand_vertex(rng, np.r_[rng.uniform(-0.5, 0.5, 3), 1.0]) for _ in range(3)])
        out = clip(t)
        assert out.shape == (3, 18)
        assert (out.view(np.uint32) == t.view(np.uint32)).all()


def test_outside_first_plane_is_dropped():
    rng = np.random.default_rng(3)
    t = np.stack([rand_vertex(rng, [-2.0 - k, 0.0, 0.0, 1.0]) for k in range(3)])
    assert clip(t).shape[0] == 0


@pytest.mark.parametrize("plane", range(6))
def test_single_plane_cases(plane):
    rng = np.random.default_rng(10 + plane)
    for _ in range(200):
        c = rng.uniform(-2.5, 2.5, (3, 4))
        c[:, 3] = rng.uniform(0.5, 2.0, 3)
        t = np.stack([rand_vertex(rng, c[k]) for k in range(3)])
        d = np.array([dist(t[k, 14:18].astype(np.float64), plane) for k in range(3)])
        # d computed in float32 as the clipper does
        d32 = np.array([dist(t[k, 14:18], plane) for k in range(3)], np.float32)
        inside = d32 >= 0
        r, out = clip_plane(plane, t)
        if inside.all():
            assert r == 3 and (out[:3].view(np.uint32) == t.view(np.uint32)).all()
            continue
        if not inside.any():
            assert r == 0
            continue
        # corner 0 of the output is an original inside corner, and the cyclic
        # order (winding) of the originals is kept
        idx = [k for k in range(3) if (out[0].view(np.uint32) == t[k].view(np.uint32)).all()]
        assert len(idx) == 1 and inside[idx[0]]
        k0 = idx[0]
        nin = inside.sum()
        assert r == (4 if nin == 2 else 3)
        new = [out[3], out[2]] if nin == 2 else [out[1], out[2]]
        if nin == 2:
            assert (out[1].view(np.uint32) == t[(k0 + 1) % 3].view(np.uint32)).all()
        for nv in new:
            # the new corners lie on the plane ...
            cp = nv[14:18].astype(np.float64)
            assert abs(dist(cp, plane)) < 1e-5 * (1 + np.abs(cp).max())
            # ... on an edge of the input, and their uv is the perspective-correct blend
            best = None
            for a in range(3):
                for b in range(3):
                    if a == b:
                        continue
                    ca, cb = t[a, 14:18].astype(np.float64), t[b, 14:18].astype(np.float64)
                    den = dist(ca, plane) - dist(cb, plane)
                    if abs(den) < 1e-12:
                        continue
                    s = dist(ca, plane) / den
                    if not (-1e-6 <= s <= 1 + 1e-6):
                        continue
                    p = ca * (1 - s) + cb * s
                    err = np.abs(p - cp).max()
                    if best is None or err < best[0]:
                        best = (err, a, b, s)
            assert best is not None and best[0] < 1e-5
            _, a, b, s = best
            wa, wb = t[a, 17], t[b, 17]
            uv = ((t[a, 3:5] * wa) * (1 - s) + (t[b, 3:5] * wb) * s) / (wa * (1 - s) + wb * s)
            assert np.allclose(nv[3:5], uv, atol=1e-5)
            for k in (5, 8, 11):
                assert abs(np.linalg.norm(nv[k:k + 3].astype(np.float64)) - 1) < 1e-5


def poly_clip_area(c):
    """Area in NDC x/y of triangle c (3 x 4 homogeneous) clipped against the
    six planes by Sutherland-Hodgman in float64."""
    poly = [np.asarray(p, np.float64) for p in c]
    for plane in range(6):
        out = []
        for i in range(len(poly)):
            a, b = poly[i], poly[(i + 1) % len(poly)]
            da, db = dist(a, plane), dist(b, plane)
            if da >= 0:
                out.append(a)
            if (da >= 0) != (db >= 0):
                s = da / (da - db)
                out.append(a * (1 - s) + b * s)
        poly = out
        if not poly:
            return 0.0
    xy = np.array([[p[0] / p[3], p[1] / p[3]] for p in poly])
    x, y = xy[:, 0], xy[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def planes_crossed(c, eps=1e-6):
    """Planes the triangle touches from both sides.  Pieces of the triangle
    split only at such planes, so with at most 3 of them the clipper's
    24-vertex list (pipeline.cpp:196-216) cannot overflow; with more, pieces
    past the list are silently dropped."""
    n = 0
    for plane in range(6):
        d = [dist(np.asarray(p, np.float64), plane) for p in c]
        n += min(d) < eps and max(d) > -eps
    return n


def test_full_clipper_area_matches_sutherland_hodgman():
    rng = np.random.default_rng(7)
    checked = dropped = 0
    for _ in range(400):
        c = rng.uniform(-2.0, 2.0, (3, 4))
        c[:, 3] = rng.uniform(0.3, 1.5, 3)
        t = np.stack([rand_vertex(rng, c[k]) for k in range(3)]).astype(np.float32)
        out = clip(t)
        assert out.shape[0] % 3 == 0 and out.shape[0] <= 24
        area = 0.0
        for j in range(0, out.shape[0], 3):
            p = out[j:j + 3, 14:18].astype(np.float64)
            xy = p[:, :2] / p[:, 3:4]
            area += 0.5 * abs((xy[1, 0] - xy[0, 0]) * (xy[2, 1] - xy[0, 1])
                              - (xy[2, 0] - xy[0, 0]) * (xy[1, 1] - xy[0, 1]))
            for q in p:
                for plane in range(6):
                    assert dist(q, plane) >= -1e-5 * (1 + np.abs(q).max())
        ref = poly_clip_area(t[:, 14:18].astype(np.float64))
        if planes_crossed(t[:, 14:18]) > 3:
            # pieces past the 24-vertex list are dropped: area can only shrink
            assert area <= ref + 1e-4 * max(1.0, ref)
            dropped += 1
        else:
            assert abs(area - ref) <= 1e-4 * max(1.0, ref), (area, ref)
            checked += out.shape[0] > 3
    assert checked > 50 and dropped > 5


def test_mat4_orders():
    # mat4 * vec4 = (c0 x + c1 y) + (c2 z + c3 w); mat4 * mat4 column by column
    rng = np.random.default_rng(5)
    a = rng.normal(size=16).astype(np.float32)
    b = rng.normal(size=16).astype(np.float32)
    v = rng.normal(size=4).astype(np.float32)
    out = np.zeros(4, np.float32)
    O.lib().orc_mat4_vec4(a.ctypes.data_as(F), v.ctypes.data_as(F), out.ctypes.data_as(F))
    A = a.reshape(4, 4)   # A[c] is column c
    want = (A[0] * v[0] + A[1] * v[1]) + (A[2] * v[2] + A[3] * v[3])
    assert (out.view(np.uint32) == want.astype(np.float32).view(np.uint32)).all()
    assert np.allclose(out, A.T.astype(np.float64) @ v, atol=1e-5)
    m = np.zeros(16, np.float32)
    O.lib().orc_mat4_mul(a.ctypes.data_as(F), b.ctypes.data_as(F), m.ctypes.data_as(F))
    assert np.allclose(m.reshape(4, 4).T, A.T.astype(np.float64) @ b.reshape(4, 4).T, atol=1e-5)
