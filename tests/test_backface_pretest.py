"""The object-space back-face pre-test of the visibility setup (k_vis, for
views whose bounding sphere lies inside the frustum).

With the clip transform's rows x, y, w written as A3 (p - e) (e: the eye in
object space, where those rows vanish), a face's screen-space cross product
is (W H / 4) det(A3) ((p0 - e) . n) / (w0 w1 w2), n = (p1 - p0) x (p2 - p0).
The pre-test culls a face when sign(det A3) ((p0 - e) . n) is below
-max(tau, 1e-3 |p0 - e|_1 |p1 - p0|_1 |p2 - p0|_1), tau = w_hi^3 / (W H
|det A3|).  These tests pin the identity on random transforms and check, on
the scenes the GPU tests use, that no face the pre-test would cull is
front-facing for the float32 screen-space test the kernels run (the GPU
parity tests then compare the pixels)."""
import numpy as np

from axiomr_amd import render as R
from axiomr_amd import scenes as S

f32 = np.float32


def eye_and_det(M):
    R3 = M[[0, 1, 3], :]
    return np.linalg.solve(R3[:, :3], -R3[:, 3]), np.linalg.det(R3[:, :3])


def test_screen_cross_identity():
    """Exact arithmetic (float64) on random projective transforms and
    triangles in front of the eye: the NDC cross product equals
    det(A3) ((p0 - e) . n) / (w0 w1 w2)."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        M = rng.normal(size=(4, 4))
        e, det = eye_and_det(M)
        P = rng.normal(size=(3, 3)) * 2.0
        C = (M @ np.c_[P, np.ones(3)].T).T
        w = C[:, 3]
        if (np.abs(w) < 1e-3).any():
            continue
        X, Y = C[:, 0] / w, C[:, 1] / w
        ndc = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0])
        n = np.cross(P[1] - P[0], P[2] - P[0])
        pred = det * np.dot(P[0] - e, n) / (w[0] * w[1] * w[2])
        assert np.isclose(ndc, pred, rtol=1e-7, atol=1e-9), (ndc, pred)


def culled_but_front(sc):
    view, proj, vp, vpm = R.camera_matrices(sc.camera_pos, sc.camera_target, sc.fov, sc.width / sc.height,
                                            vp_w=sc.width, vp_h=sc.height)
    M = np.array(vp, np.float64).reshape(4, 4).T @ np.array(sc.model, np.float64).reshape(4, 4).T
    M32 = M.astype(f32)
    P = sc.vertices[:, :3].astype(f32)
    C = ((M32[:, 0] * P[:, 0:1] + M32[:, 1] * P[:, 1:2]) + (M32[:, 2] * P[:, 2:3] + M32[:, 3])).astype(f32)
    w = C[:, 3]
    assert (w > 0).all()
    inv = (f32(1) / w).astype(f32)
    sx = ((((C[:, 0] * inv) + f32(1)) * f32(0.5)) * f32(sc.width)).astype(f32)
    sy = ((((C[:, 1] * inv) + f32(1)) * f32(0.5)) * f32(sc.height)).astype(f32)
    F = sc.faces.astype(np.int64)
    dx1, dy1 = sx[F[:, 1]] - sx[F[:, 0]], sy[F[:, 1]] - sy[F[:, 0]]
    dx2, dy2 = sx[F[:, 2]] - sx[F[:, 0]], sy[F[:, 2]] - sy[F[:, 0]]
    back = (dx1 * dy2 - dx2 * dy1) < 0
    e, det = eye_and_det(M)
    cen = P.astype(np.float64).mean(0)
    rad = np.linalg.norm(P.astype(np.float64) - cen, axis=1).max()
    w_hi = M[3, :3] @ cen + M[3, 3] + np.linalg.norm(M[3, :3]) * rad
    tau = f32(w_hi ** 3 / (sc.width * sc.height * abs(det)))
    e32 = e.astype(f32)
    p0, p1, p2 = P[F[:, 0]], P[F[:, 1]], P[F[:, 2]]
    a, b, c = p0 - e32, p1 - p0, p2 - p0
    n = np.cross(b, c).astype(f32)
    d = f32(np.sign(det)) * ((a[:, 0] * n[:, 0] + a[:, 1] * n[:, 1]) + a[:, 2] * n[:, 2])
    la = np.abs(a).sum(1, dtype=f32) + np.abs(e32).sum(dtype=f32)
    lim = np.maximum(tau, f32(1e-3) * la * np.abs(b).sum(1, dtype=f32) * np.abs(c).sum(1, dtype=f32))
    pre = d < -lim
    return int((pre & ~back).sum()), float(pre.mean()), float(back.mean())


def test_pretest_never_culls_a_front_face():
    scenes = [S.icosphere_scene(n=64, width=1920, height=1080, tex=16),
              S.icosphere_scene(n=32, width=320, height=240, tex=16),
              S.uv_sphere_scene(),
              S.pbr_scene(n=24, width=160, height=120, tex=16)]
    scenes += [S.icosphere_scene(n=48, width=640, height=360, tex=16, yaw=y, pitch=30.0) for y in (0, 45, 90, 170)]
    for sc in scenes:
        wrong, pre, back = culled_but_front(sc)
        assert wrong == 0, sc.name
        assert pre > 0.5 * back, (sc.name, pre, back)     # it does cull most back faces
