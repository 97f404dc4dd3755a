"""Host-side helpers of the benchmark and the autotuner, without a GPU: the
traffic lookup for the two-launch resolve, the overflow-tolerant warm-up and
the two-round coordinate descent."""
import json
import sys
from pathlib import Path

import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import bench  # noqa: E402
from axiomr_amd import autotune as AT  # noqa: E402
from axiomr_amd._native import AXR_ERR_INVALID, AXR_ERR_OVERFLOW, AxrError  # noqa: E402

TUN = {"bin_capacity": 128, "raster_mode": 1, "raster_threads": 128, "small_footprint": 128, "tile_size": 16}


def _traffic_file(tmp_path, kernels, tunables=TUN):
    f = tmp_path / "traffic.json"
    f.write_text(json.dumps({"tunables": tunables,
                             "kernels": {k: {"traffic_bytes": v} for k, v in kernels.items()}}))
    return f


def test_traffic_of_a_launch_pair(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "TRAFFIC_FILE", _traffic_file(tmp_path, {
        "k_shade<16,4>": 42_000_000.4, "k_finish4<16>": 22_000_000.2, "k_finish<16>": 1.0, "k_vis<16,false>": 5.0}))
    assert bench.measured_traffic("k_shade+k_finish4", TUN) == 64_000_001
    assert bench.measured_traffic("k_vis", TUN) == 5
    # a part that is missing, or tunables that differ: no figure
    assert bench.measured_traffic("k_shade+k_resolve", TUN) is None
    assert bench.measured_traffic("k_shade+k_finish4", {**TUN, "tile_size": 8}) is None


def test_traffic_of_one_instantiation(tmp_path, monkeypatch):
    # the batch's kernels and the plain frame's share names; the template
    # argument picks one
    f = _traffic_file(tmp_path, {"k_raster<32,128,true,4>": 7.0, "k_raster<32,128,false,4>": 3.0,
                                 "k_shade<16,4,false>": 2.0, "k_shade<16,4,true>": 9.0})
    monkeypatch.setattr(bench, "TRAFFIC_FILE", f)
    assert bench.measured_traffic("k_raster", TUN) is None
    assert bench.measured_traffic("k_raster", TUN, instance="true") == 7
    assert bench.measured_traffic("k_shade", TUN, instance="false") == 2
    assert bench.measured_traffic("k_raster", TUN, path=f, instance="false") == 3
    assert bench.measured_traffic("k_raster", TUN, path=tmp_path / "absent.json", instance="true") is None


def test_traffic_needs_one_match_per_part(tmp_path, monkeypatch):
    # "k_finish<" must not also match "k_finish4<"; two instantiations of one
    # name are ambiguous
    monkeypatch.setattr(bench, "TRAFFIC_FILE", _traffic_file(tmp_path, {
        "k_finish<16>": 3.0, "k_finish4<16>": 4.0, "k_vis<8,false>": 1.0, "k_vis<16,false>": 2.0}))
    assert bench.measured_traffic("k_finish", TUN) == 3
    assert bench.measured_traffic("k_vis", TUN) is None


class _FakeCtx:
    """sync() raises the queued errors one by one, then succeeds."""

    def __init__(self, errors):
        self.errors = list(errors)
        self.syncs = 0

    def sync(self):
        self.syncs += 1
        if self.errors:
            raise self.errors.pop(0)


def test_warm_reruns_after_an_overflow():
    calls = []
    ctx = _FakeCtx([AxrError(AXR_ERR_OVERFLOW, "grown"), AxrError(AXR_ERR_OVERFLOW, "grown")])
    bench.warm(ctx, lambda: calls.append(1))
    assert len(calls) == 3 and ctx.syncs == 3


def test_warm_passes_other_errors_and_gives_up():
    with pytest.raises(AxrError) as e:
        bench.warm(_FakeCtx([AxrError(AXR_ERR_INVALID, "bad")]), lambda: None)
    assert e.value.code == AXR_ERR_INVALID
    with pytest.raises(AxrError) as e:
        bench.warm(_FakeCtx([AxrError(AXR_ERR_OVERFLOW, "again")] * 3), lambda: None, tries=3)
    assert e.value.code == AXR_ERR_OVERFLOW


class _TuneCtx:
    def __init__(self, cfg):
        self.cfg = dict(cfg)

    def config(self):
        return dict(self.cfg)

    def set_config(self, **kw):
        self.cfg.update(kw)

    def sync(self):
        pass


def test_two_rounds_of_descent(monkeypatch):
    """A cost whose best launch shape depends on the footprint: the first
    round picks the footprint under the starting shape, the second round
    finds the shape that is best under that footprint.  No point is timed
    twice."""
    def cost(c):
        shape = (c["tile_size"], c["raster_threads"])
        t = 10.0 + abs(c["small_footprint"] - 64) / 64.0
        if shape == (16, 128):
            t -= 0.5                  # beats the start under every footprint
        if shape == (32, 64) and c["small_footprint"] == 64:
            t -= 5.0                  # best overall, seen only under footprint 64
        return t + (0.0 if c["raster_mode"] == 1 else 3.0)

    ctx = _TuneCtx({"tile_size": 8, "raster_threads": 256, "bin_capacity": 128, "small_footprint": 12,
                    "raster_mode": 1})
    monkeypatch.setattr(AT, "time_config", lambda c, frame, reps, inner: cost(c.config()))
    res = AT.tune(ctx, lambda: (lambda: None))
    best = res["best"]
    assert (best["tile_size"], best["raster_threads"], best["small_footprint"], best["raster_mode"]) == (32, 64, 64, 1)
    assert res["seconds"] == pytest.approx(5.0)
    keys = [tuple(sorted((k, v) for k, v in t.items() if k != "seconds")) for t in res["trials"]]
    assert len(keys) == len(set(keys))
    assert ctx.config() == best


def test_capacity_takes_the_smallest_within_slack(monkeypatch):
    """The bin capacity sizes the pair regions, so a larger one is kept only
    when it is more than 1 % faster: here 1024 is 0.5 % faster than 256
    (kept: 256) and 256 is 2 % faster than 128 (so not 128)."""
    cap_cost = {64: 10.5, 128: 10.2, 256: 10.0, 512: 10.0, 1024: 9.95}

    def cost(c):
        t = cap_cost[c["bin_capacity"]] if c["raster_mode"] == 0 else 20.0
        return t + (0.0 if (c["tile_size"], c["raster_threads"], c["small_footprint"]) == (32, 128, 256) else 1.0)

    ctx = _TuneCtx({"tile_size": 32, "raster_threads": 128, "bin_capacity": 128, "small_footprint": 256,
                    "raster_mode": 0})
    monkeypatch.setattr(AT, "time_config", lambda c, frame, reps, inner: cost(c.config()))
    res = AT.tune(ctx, lambda: (lambda: None))
    assert res["best"]["bin_capacity"] == 256, res["best"]
    assert res["seconds"] == pytest.approx(10.0)
    assert ctx.config() == res["best"]
