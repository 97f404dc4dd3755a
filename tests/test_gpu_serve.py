"""The served form: requests admitted continuously and drawn in batches on
several lanes must give, request by request, exactly what the plain
per-frame path gives for the same request alone (and what the CPU checker
gives), whatever else shares their batch; admission under a tight HBM budget
defers requests without changing their pixels."""
import numpy as np
import pytest

import harness as H
from axiomr_amd import scenes as S
from axiomr_amd.resident import GpuServeScene
from axiomr_amd.serve import DONE, Server
from axiomr_amd._native import AxrError, AXR_ERR_BUDGET

pytestmark = pytest.mark.gpu

SMALL_RES = ((256, 256, 0.5), (320, 180, 0.35), (480, 270, 0.15))


@pytest.fixture(scope="module")
def served(ctx):
    ctx.set_config(tile_size=16, raster_threads=256, bin_capacity=128, small_footprint=32)
    wl = S.serve_workload(nmeshes=6, fmin=2_000, fmax=200_000, tex=256, seed=41)
    gs = GpuServeScene(ctx, wl)
    yield wl, gs
    gs.close()


def _same(a, b, tag):
    for x, y, name in zip(a, b, ("colour", "depth", "ids")):
        if x is None or y is None:
            continue
        assert x.shape == y.shape, f"{tag}: {name} shape"
        assert (x.view(np.uint8) == y.view(np.uint8)).all(), f"{tag}: {name} differs from the request drawn alone"


def _requests(wl, n, seed, res=SMALL_RES):
    return wl.requests(n, seed, res)


def test_mix_equals_alone(ctx, served):
    wl, gs = served
    views = _requests(wl, 48, 1)
    with Server(ctx, lanes=3, max_batch_views=8, with_ids=True) as srv:
        srv.pause(True)           # the whole stream queues, then is admitted at once
        tickets = [srv.submit(gs.meshes[v.mesh], v.width, v.height, gs.params(v)) for v in views]
        srv.pause(False)
        for t in tickets:
            srv.wait(t, 60.0)
            assert srv.poll(t) == DONE
        results = [srv.read(t) for t in tickets]
        st = srv.stats()
        for t in tickets:
            srv.release(t)
    assert st["completed"] == 48 and st["failed"] == 0
    assert st["batches"] <= 48 // 8 + 3, st     # requests shared batches
    assert st["max_views_in_batch"] == 8
    assert st["max_lanes_busy"] == 3 and st["launches_overlapped"] >= 2, st
    covered = 0
    for i, v in enumerate(views):
        alone = gs.draw_plain(v, with_ids=True)
        _same(results[i], alone, f"request {i} ({v.width}x{v.height}, mesh {v.mesh})")
        covered += int(np.isfinite(alone[1]).sum())
    assert covered > 100_000


def test_served_requests_match_checker(ctx, served):
    wl, gs = served
    views = _requests(wl, 6, 2)
    with Server(ctx, lanes=2, with_ids=True) as srv:
        tickets = [srv.submit(gs.meshes[v.mesh], v.width, v.height, gs.params(v)) for v in views]
        for t, v in zip(tickets, views):
            srv.wait(t, 60.0)
            rep = H.compare(srv.read(t), H.render_oracle(wl.scene(v)))
            H.assert_parity(rep, f"served {v.width}x{v.height} mesh {v.mesh}")
            srv.release(t)


def test_full_resolution_mix(ctx, served):
    """The 512^2 / 1280x720 / 1920x1080 mix, several lanes, against alone."""
    wl, gs = served
    views = _requests(wl, 24, 3, S.SERVE_RESOLUTIONS)
    with Server(ctx, lanes=4, max_batch_views=6) as srv:
        srv.pause(True)
        tickets = [srv.submit(gs.meshes[v.mesh], v.width, v.height, gs.params(v)) for v in views]
        srv.pause(False)
        for t in tickets:
            srv.wait(t, 60.0)
        results = [srv.read(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == 24
    for i, v in enumerate(views):
        _same(results[i][:2], gs.draw_plain(v)[:2], f"request {i} ({v.width}x{v.height})")


def test_budget_defers_then_admits(ctx, served):
    wl, gs = served
    views = [S.ViewSpec(mesh=i % len(wl.meshes), width=480, height=270,
                        camera_pos=S.orbit_position(3.0, 40.0 * i, 10.0)) for i in range(8)]
    small = dict(lanes=2, max_batch_views=4, max_batch_faces=1 << 20, max_batch_tiles=1 << 14)
    with Server(ctx, **small) as probe:
        fixed = probe.stats()["fixed_bytes"]
    block = 2 * 480 * 270 * 4 * 6 // 5          # one 480x270 colour+depth block, class slack included
    budget = fixed + 2 * block + block // 2      # room for two results at a time
    with Server(ctx, hbm_budget=budget, **small) as srv:
        with pytest.raises(AxrError) as e:
            srv.submit(gs.meshes[0], 2048, 2048, gs.params(views[0]))
        assert e.value.code == AXR_ERR_BUDGET
        tickets = [srv.submit(gs.meshes[v.mesh], v.width, v.height, gs.params(v)) for v in views]
        srv.wait(tickets[0], 60.0)
        srv.wait(tickets[1], 60.0)
        # nothing else fits until a result is released
        st = srv.stats()
        assert st["completed"] == 2 and st["fb_peak_bytes"] <= budget - fixed
        for i, t in enumerate(tickets):
            srv.wait(t, 60.0)
            _same(srv.read(t), gs.draw_plain(views[i]), f"deferred request {i}")
            srv.release(t)
        st = srv.stats()
    assert st["completed"] == 8 and st["deferred"] > 0 and st["rejected"] == 1
    assert st["fb_peak_bytes"] <= budget - fixed


def test_replay_open_loop(ctx, served):
    wl, gs = served
    views = _requests(wl, 300, 4)
    reqs = [gs.request(v) for v in views]
    arrivals = S.poisson_arrivals(len(reqs), 3000.0, seed=4)
    with Server(ctx, lanes=3) as srv:
        lat = srv.replay(reqs, arrivals)
        st = srv.stats()
    assert (lat >= 0).all() and np.isfinite(lat).all()
    assert st["completed"] == 300 and st["fb_in_use_bytes"] == 0
    assert float(np.percentile(lat, 50)) < 0.5


def test_lane_overflow_is_redrawn():
    """A lane sized for small batches (max_batch_faces 4096) takes a request
    with far more large and clipped triangles than its arena holds: the
    batch's kernels flag the overflow, the lane grows its arena to the
    reported needs and draws the request again, and its pixels then match
    the checker's."""
    from axiomr_amd import _native as N
    from axiomr_amd import render as R
    sc = S.soup_scene(ntri=60_000, width=480, height=270, seed=9)
    ref = H.render_oracle(sc)
    c = R.Context(0)
    try:
        c.set_config(tile_size=16, raster_mode=N.RASTER_VISIBILITY, small_footprint=12)
        gs = H.GpuScene(c, sc)
        try:
            with Server(c, lanes=1, max_batch_views=4, max_batch_faces=4096, with_ids=True) as srv:
                t = srv.submit(gs.mesh, sc.width, sc.height, H.params_of(sc))
                srv.wait(t, 60.0)
                assert srv.poll(t) == DONE
                got = srv.read(t)
                st = srv.stats()
                srv.release(t)
        finally:
            gs.close()
    finally:
        c.close()
    assert st["overflow_retries"] >= 1, st
    assert st["failed"] == 0 and st["completed"] == 1, st
    H.assert_parity(H.compare(got, ref), "served soup after an arena overflow")
