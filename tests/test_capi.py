"""The C ABI library: loads without a GPU, exports every entry point that
include/axiomr.h declares, and its host-only helpers (scene generators,
camera) produce what they document."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import harness as H
from axiomr_amd import _native as N
from axiomr_amd import scenes as S

HEADER = Path(__file__).resolve().parent.parent / "include" / "axiomr.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("axr_ctx_create", "axr_draw", "axr_mesh_create", "axr_texture_create", "axr_fb_create",
                 "axr_fb_clear", "axr_fb_read", "axr_batch_create", "axr_batch_launch", "axr_sync"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(N.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_the_header():
    assert set(declared_functions()) <= set(N.SIGNATURES)


def test_version_and_device_query():
    assert b"gfx950" in N.lib().axr_version()
    n = C.c_int(-1)
    assert N.lib().axr_device_count(C.byref(n)) == 0
    assert n.value >= 0


def test_context_without_gpu_reports_no_device():
    n = C.c_int(0)
    N.lib().axr_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert N.lib().axr_ctx_create(0, 0, C.byref(h)) == N.AXR_ERR_NO_DEVICE
    assert not h.value


def test_null_handles_are_rejected():
    L = N.lib()
    assert L.axr_ctx_destroy(None) == N.AXR_ERR_INVALID
    assert L.axr_sync(None) == N.AXR_ERR_INVALID
    assert L.axr_draw(None, None, None, None) == N.AXR_ERR_INVALID
    assert L.axr_fb_clear(None, 0, 0.0) == N.AXR_ERR_INVALID
    assert L.axr_last_error(None) == b"null context"


def test_uv_sphere_counts_and_unit_normals():
    v, f = S.uv_sphere(100, 100)
    assert v.shape == (101 * 101, 14) and f.shape == (20000, 3)
    assert f.max() < v.shape[0]
    r = np.linalg.norm(v[:, :3], axis=1)
    assert np.allclose(r, 1.0, atol=1e-6)
    assert np.allclose(np.linalg.norm(v[:, 5:8], axis=1), 1.0, atol=1e-6)
    # outward, counter-clockwise from outside (non-degenerate faces)
    p = v[:, :3].astype(np.float64)
    cr = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    cen = p[f].mean(axis=1)
    big = np.linalg.norm(cr, axis=1) > 1e-9
    assert (np.einsum("ij,ij->i", cr[big], cen[big]) > 0).all()
    assert (~big).sum() == 200   # the pole rows are degenerate


@pytest.mark.parametrize("n", [1, 3, 8])
def test_icosphere_counts_and_displacement(n):
    v, f = S.icosphere(n, 0.1, 1)
    assert v.shape == (10 * n * n + 2, 14) and f.shape == (20 * n * n, 3)
    r = np.linalg.norm(v[:, :3], axis=1)
    assert r.min() >= 0.9 - 1e-6 and r.max() <= 1.1 + 1e-6
    p = v[:, :3].astype(np.float64)
    cr = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    assert (np.einsum("ij,ij->i", cr, p[f].mean(axis=1)) > 0).all()
    # every edge shared by exactly two faces: a closed manifold
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()
    assert np.allclose(np.linalg.norm(v[:, 5:8], axis=1), 1.0, atol=1e-5)


def test_soup_statistics():
    v, f, c = S.soup(20000, 1920, 1080, 60.0, 2.0, 1.5, 0.1, 3)
    assert v.shape == (60000, 14) and f.shape == (20000, 3)
    tri = v[:, :3].reshape(-1, 3, 3)
    # one view depth per triangle in [0.1, 100]
    d = -tri[:, :, 2]
    assert np.allclose(d.min(axis=1), d.max(axis=1))
    assert d.min() >= 0.1 - 1e-6 and d.max() <= 100.0 + 1e-4
    # exact duplicates of earlier triangles: about 10%
    _, first = np.unique(tri.reshape(len(tri), -1), axis=0, return_index=True)
    dup = 1.0 - len(first) / len(tri)
    assert 0.08 < dup < 0.12
    # screen areas: log-normal with median ~2 px
    t = np.tan(math.radians(30.0))
    sx = (tri[:, :, 0] / (d * t * 1920 / 1080) + 1) * 960
    sy = (tri[:, :, 1] / (d * t) + 1) * 540
    area = 0.5 * ((sx[:, 1] - sx[:, 0]) * (sy[:, 2] - sy[:, 0]) - (sx[:, 2] - sx[:, 0]) * (sy[:, 1] - sy[:, 0]))
    assert (area > 0).all()                       # counter-clockwise on screen
    assert 1.7 < np.median(area) < 2.3


def test_textures():
    ch = S.checker(256, 256, 8)
    assert (ch[0, 0] != ch[0, 8]).any() and (ch[0, 0] == ch[8, 8]).all()
    nm = S.normal_map(64, 64, 0.7, 6).astype(np.float64)
    n = nm[..., :3] / 255.0 * 2.0 - 1.0
    assert np.allclose(np.linalg.norm(n, axis=-1), 1.0, atol=0.02)
    assert n[..., 2].min() >= 0.7 - 0.01
    nz = S.noise_texture(128, 128, 5)
    assert nz[..., 3].min() == 255 and 100 < nz[..., :3].mean() < 155


def test_camera_matches_closed_form():
    view, proj, vp, vpm = H.R.camera_matrices((0, 0, 3), (0, 0, 0), 60.0, 4.0 / 3.0, 0.1, 100.0, 800, 600)
    V = view.reshape(4, 4).T.astype(np.float64)
    P = proj.reshape(4, 4).T.astype(np.float64)
    expect_v = np.eye(4)
    expect_v[2, 3] = -3.0
    assert np.allclose(V, expect_v, atol=1e-6)
    f = 1.0 / math.tan(math.radians(30.0))
    assert abs(P[0, 0] - f / (4.0 / 3.0)) < 1e-5 and abs(P[1, 1] - f) < 1e-5
    assert abs(P[2, 2] + 100.1 / 99.9) < 1e-5 and abs(P[2, 3] + 20.0 / 99.9) < 1e-5 and P[3, 2] == -1.0
    assert np.allclose(vp.reshape(4, 4).T, P @ V, atol=1e-5)
    M = vpm.reshape(4, 4).T
    assert M[0, 0] == 400 and M[1, 1] == -300 and M[0, 3] == 400 and M[1, 3] == 300


def test_orbit_camera_looks_at_target():
    pos = S.orbit_position(3.0, 25.0, 15.0)
    view, proj, vp, _ = H.R.camera_matrices(pos, (0, 0, 0), 60.0, 16 / 9)
    c = vp.reshape(4, 4).T.astype(np.float64) @ np.array([0, 0, 0, 1.0])
    assert abs(c[0] / c[3]) < 1e-5 and abs(c[1] / c[3]) < 1e-5   # target at the screen centre


def test_view_limits():
    """A view submits fewer than 2^29 faces (the 32-bit tie-break key is
    rank * 8 + clip sub-triangle) on a framebuffer of 1..32768 pixels per
    edge; axr_draw, batch creation and server submission apply this check."""
    lib = N.lib()
    assert lib.axr_check_view(0, 1, 1) == N.AXR_OK
    assert lib.axr_check_view((1 << 29) - 1, 32768, 32768) == N.AXR_OK
    # the largest key a view can produce still fits 32 bits
    assert ((1 << 29) - 1) * 8 + 7 == 0xFFFFFFFF
    for faces, w, h in (((1 << 29), 64, 64), ((1 << 32), 64, 64), ((1 << 63), 64, 64), (10, 0, 64), (10, 64, 0),
                        (10, 32769, 64), (10, 64, 32769), (10, -1, 64)):
        assert lib.axr_check_view(faces, w, h) == N.AXR_ERR_INVALID, (faces, w, h)
