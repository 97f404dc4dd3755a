"""The CPU checker against the exact-arithmetic known answers in
tests/golden/raster_kat.json (written by tests/golden/make_golden.py)."""
import json
from pathlib import Path

import numpy as np

import harness as H
from test_oracle_raster import identity_params, ndc_scene

KAT = Path(__file__).resolve().parent / "golden" / "raster_kat.json"


def load_kat():
    d = json.loads(KAT.read_text())
    tris = np.array(d["triangles_ndc"], np.float32)
    ids = np.array(d["expected_ids"], np.uint32)
    depth = np.array([[np.nan if v is None else v for v in row] for row in d["expected_depth"]], np.float64)
    return d["width"], d["height"], tris, ids, depth


def check_against_kat(render):
    W, Hh, tris, ids, depth = load_kat()
    scene = ndc_scene(tris, W, Hh)
    bgra, d, got = render(scene, [identity_params()])
    assert (got == ids).all(), f"{(got != ids).sum()} pixels with another winner"
    cov = ids != H.EMPTY_ID
    assert np.isinf(d[~cov]).all()
    assert np.abs(d[cov] - depth[cov]).max() < 1e-5
    return cov.sum()


def test_oracle_matches_known_answers():
    assert check_against_kat(lambda sc, draws: H.render_oracle(sc, draws=draws)) > 1000


def test_fixture_is_consistent():
    W, Hh, tris, ids, depth = load_kat()
    assert ids.shape == (Hh, W) and tris.shape[1:] == (3, 3)
    # the culled back face, the zero-area and the later duplicate never win
    won = set((ids[ids != H.EMPTY_ID] >> 3).tolist())
    assert not won & {3, 4, 6}
