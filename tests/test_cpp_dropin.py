"""The C++ mirror of the reference interface (axiomr_amd/cpp/axiomr.hpp):
the example frame loop builds with plain g++ against the C ABI, fails
loudly without a GPU, and on the GPU draws exactly what the Python path
draws for the same scene."""
import subprocess

import numpy as np
import pytest

import harness as H
from axiomr_amd import build as B
from axiomr_amd import scenes as S


def test_example_builds_and_needs_a_gpu():
    exe = B.build_example()
    import ctypes as C
    from axiomr_amd import _native as N
    n = C.c_int(0)
    N.lib().axr_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is present")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "error -6" in r.stderr


@pytest.mark.gpu
def test_example_matches_python_path(ctx):
    exe = B.EXAMPLE_BIN
    assert exe.exists(), "build the library first"
    r = subprocess.run([str(exe), "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    covered, _, hexhash = r.stdout.split()[1:4]
    bgra, depth, _ = H.render_gpu(ctx, S.uv_sphere_scene(), tile=16, with_ids=False)
    assert int(covered) == int(np.isfinite(depth).sum())
    # hash the way the example does (BGRA bytes, then depth bytes)
    h = 1469598103934665603
    data = np.concatenate([bgra.reshape(-1), depth.view(np.uint8).reshape(-1)])
    for x in data.tobytes():
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    assert f"{h:016x}" == hexhash


@pytest.mark.gpu
def test_example_draws_an_obj_like_the_python_path(ctx, tmp_path):
    """The reference Renderer's flow (Mesh(path) with its .mtl and textures,
    then drawMesh) through the C++ mirror, against the Python mirror."""
    PIL = pytest.importorskip("PIL.Image")
    from test_gpu_ingest import write_assets
    from axiomr_amd import render as R
    path = write_assets(tmp_path)
    png = tmp_path / "frame.png"
    r = subprocess.run([str(B.EXAMPLE_BIN), "--obj", path, str(png)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    covered, _, hexhash = r.stdout.split()[1:4]
    mesh = R.Mesh.from_obj(ctx, path)
    fb = R.Framebuffer(ctx, 800, 600)
    cam = R.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 60.0, 800 / 600)
    cam.setViewport(0, 0, 800, 600)
    cam.update(0.0)
    pipe = R.TiledPipeline(cam, fb)
    pipe.setShader(R.PhongShader(lightDirection=(0.0, -1.0, 0.0)))
    try:
        fb.clearColor(R.Color(0, 0, 0, 255))
        fb.clearDepth()
        pipe.drawMesh(np.eye(4, dtype=np.float32).reshape(16), mesh)
        bgra, depth, _ = fb.read()
    finally:
        fb.close()
        mesh.close()
    assert int(covered) == int(np.isfinite(depth).sum()) > 10_000
    h = 1469598103934665603
    data = np.concatenate([bgra.reshape(-1), depth.view(np.uint8).reshape(-1)])
    for x in data.tobytes():
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    assert f"{h:016x}" == hexhash
    img = np.asarray(PIL.open(str(png)).convert("RGBA"))
    np.testing.assert_array_equal(img, bgra[::-1][..., [2, 1, 0, 3]])
