"""GPU path vs the CPU checker on the same seeded scenes: coverage, depth and
winning primitive bit-exact, RGBA8 colour within 1 LSB per channel."""
import numpy as np
import pytest

import harness as H
from axiomr_amd import scenes as S

pytestmark = pytest.mark.gpu


def _run(ctx, scene, tile=16, draws=None):
    gpu = H.render_gpu(ctx, scene, draws=draws, tile=tile)
    ref = H.render_oracle(scene, draws=draws)
    rep = H.compare(gpu, ref)
    print(f"\n{scene.name} tile={tile}: {rep}", flush=True)
    H.assert_parity(rep, scene.name)
    return rep


@pytest.mark.parametrize("tile", [16, 8, 32])
def test_uv_sphere_800x600(ctx, tile):
    rep = _run(ctx, S.uv_sphere_scene(), tile=tile)
    assert rep["covered"] > 100_000


def test_icosphere_small(ctx):
    sc = S.icosphere_scene(n=40, width=640, height=360, tex=256)
    rep = _run(ctx, sc)
    assert rep["covered"] > 10_000


def test_icosphere_partial_tiles(ctx):
    # widths and heights that are not multiples of 16 exercise partial bins
    sc = S.icosphere_scene(n=24, width=333, height=197, tex=128, yaw=70.0, pitch=-20.0)
    for tile in (8, 16, 32):
        _run(ctx, sc, tile=tile)


def test_soup_with_ties_and_clipping(ctx):
    sc = S.soup_scene(ntri=60_000, width=480, height=270, seed=9)
    rep = _run(ctx, sc)
    assert rep["covered"] > 50_000
