"""GPU path on the hand-placed corner cases: the exact-arithmetic known
answers, the per-pixel brute force scenes, -0/+0 depth ties, duplicates and
degenerate triangles, for every bin size."""
import numpy as np
import pytest

import harness as H
import rastercheck as RC
from test_golden import load_kat
from test_oracle_raster import identity_params, mat_mul, ndc_scene, random_tris, transform

pytestmark = pytest.mark.gpu
f32 = np.float32


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_known_answers(ctx, tile):
    W, Hh, tris, ids, depth = load_kat()
    scene = ndc_scene(tris, W, Hh)
    _, d, got = H.render_gpu(ctx, scene, draws=[identity_params()], tile=tile)
    assert (got == ids).all()
    cov = ids != H.EMPTY_ID
    assert np.isinf(d[~cov]).all() and np.abs(d[cov] - depth[cov]).max() < 1e-5


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("W,Hh,T,scale,seed", [(64, 48, 60, 0.15, 1), (72, 40, 200, 0.05, 2), (37, 53, 120, 0.3, 3),
                                               (160, 90, 400, 0.04, 4)])
def test_brute_force_scenes(ctx, tile, W, Hh, T, scale, seed):
    rng = np.random.default_rng(seed)
    tris = random_tris(rng, T, scale)
    scene = ndc_scene(tris, W, Hh)
    gpu = H.render_gpu(ctx, scene, draws=[identity_params()], tile=tile)
    clip = transform(np.eye(4, dtype=f32).reshape(16), tris)
    sx, sy, z = RC.screen_setup(clip, W, Hh)
    keep = ~RC.backface(sx, sy)
    depth, ids = RC.resolve(sx[keep], sy[keep], z[keep], (np.arange(T, dtype=np.uint32) * 8)[keep], W, Hh)
    assert (gpu[1].view(np.uint32) == depth.view(np.uint32)).all()
    assert (gpu[2] == ids).all()
    rep = H.compare(gpu, H.render_oracle(scene, draws=[identity_params()]))
    H.assert_parity(rep, f"brute {W}x{Hh} tile {tile}")


def test_signed_zero_ties(ctx):
    vp = np.eye(4, dtype=f32)
    vp[0, 2] = vp[1, 2] = vp[3, 2] = -0.0
    vp = vp.reshape(16)
    model = np.eye(4, dtype=f32)
    model[3, 2] = -0.0
    model = model.reshape(16)
    t = np.array([[[-0.9, -0.9, 0.0], [-0.05, -0.9, 0.0], [-0.5, -0.05, 0.0]]], f32)
    neg = t.copy()
    neg[..., 2] = -0.0
    for tris in (np.concatenate([neg, t]), np.concatenate([t, neg])):
        scene = ndc_scene(tris, 48, 48)
        gpu = H.render_gpu(ctx, scene, draws=[identity_params(model, vp)])
        ref = H.render_oracle(scene, draws=[identity_params(model, vp)])
        H.assert_parity(H.compare(gpu, ref), "signed zero")
        cov = gpu[2] != H.EMPTY_ID
        assert cov.sum() > 50 and (np.signbit(gpu[1][cov]) == np.signbit(tris[0, 0, 2])).all()


def test_duplicates_and_degenerates(ctx):
    rng = np.random.default_rng(11)
    base = random_tris(rng, 40, 0.2)
    degenerate = np.array([
        [[-0.5, -0.5, 0.1], [0.5, 0.5, 0.1], [0.0, 0.0, 0.1]],
        [[0.1, 0.1, 0.2], [0.1, 0.1, 0.2], [0.1, 0.1, 0.2]],
        [[-0.3, -0.3, 0.3], [-0.3 + 1e-7, -0.3, 0.3], [-0.3, -0.3 + 1e-7, 0.3]]], f32)
    tris = np.concatenate([base, degenerate, base[::-1], base[:10]])
    scene = ndc_scene(tris, 64, 64)
    for tile in (8, 16, 32):
        gpu = H.render_gpu(ctx, scene, draws=[identity_params()], tile=tile)
        H.assert_parity(H.compare(gpu, H.render_oracle(scene, draws=[identity_params()])), f"dups tile {tile}")
        hit = gpu[2][gpu[2] != H.EMPTY_ID] >> 3
        assert (hit < 40).all()
