"""The autotuner's search, workload keys and cache (empm_amd.autotune),
without a GPU: the timings come from a cost model of the configuration."""
import json

import pytest

from empm_amd import autotune as at


def cost(cfg):
    """Separable, with its minimum at scatter_wg 128, scatter_per_cu 2,
    heavy_chunks 4, dynamic_rounds 8, node_grid_cap 8192."""
    want = {"scatter_wg": 128, "scatter_per_cu": 2, "heavy_chunks": 4, "dynamic_rounds": 8, "node_grid_cap": 8192}
    return 1.0 + sum(0.1 * (cfg[k] != v) for k, v in want.items())


def test_coordinate_search_finds_the_separable_minimum():
    calls = []
    best, t, trace = at.coordinate_search(lambda c: calls.append(c) or cost(c))
    assert best == {"scatter_wg": 128, "scatter_per_cu": 2, "heavy_chunks": 4, "dynamic_rounds": 8,
                    "node_grid_cap": 8192}
    assert t == pytest.approx(1.0)
    assert len(trace) == len(calls)
    # the first timing is the default configuration
    assert trace[0][0] == at.DEFAULT
    # every configuration timed is inside the space
    for c, _ in trace:
        for k, v in c.items():
            assert v in at.SPACE[k]


def test_gains_below_the_tolerance_keep_the_current_value():
    def flat(c):
        return 1.0 - (0.005 if c["node_grid_cap"] == 16384 else 0.0)
    best, _, trace = at.coordinate_search(flat, rel_tol=0.01)
    assert best == at.DEFAULT
    # a pass without a change ends the search: one pass over every other value
    assert len(trace) == 1 + sum(len(v) - 1 for v in at.SPACE.values())


def test_coupled_parameters_need_a_second_pass():
    # node_grid_cap 1024 only pays once scatter_wg is 512 (searched before it)
    def coupled(c):
        t = 1.0 - 0.05 * (c["scatter_wg"] == 512)
        if c["scatter_wg"] == 512 and c["node_grid_cap"] == 1024:
            t -= 0.05
        return t
    best, t, _ = at.coordinate_search(coupled)
    assert best["scatter_wg"] == 512 and best["node_grid_cap"] == 1024
    assert t == pytest.approx(0.9)


class FakeCtx:
    def __init__(self):
        self.tuning = dict(at.DEFAULT)
        self.sets = 0

    def set_tuning(self, t):
        self.tuning = dict(t)
        self.sets += 1

    def get_tuning(self):
        return dict(self.tuning)


def test_tune_sets_the_best_and_caches_it(tmp_path, monkeypatch):
    monkeypatch.setenv("EMPM_TUNING_CACHE", str(tmp_path / "tuning.json"))
    ctx = FakeCtx()
    timer = lambda run, reps: cost(ctx.tuning)
    r = at.tune(ctx, lambda: None, key="k1", timer=timer)
    assert not r["cached"] and ctx.tuning["scatter_wg"] == 128 and ctx.tuning["node_grid_cap"] == 8192
    assert r["default_seconds"] == pytest.approx(cost(at.DEFAULT))
    saved = json.load(open(tmp_path / "tuning.json"))
    assert saved["k1"]["tuning"] == ctx.tuning
    # a second tuner finds it in the file and times nothing
    ctx2 = FakeCtx()
    r2 = at.tune(ctx2, lambda: None, key="k1", timer=lambda run, reps: pytest.fail("timed a cached workload"))
    assert r2["cached"] and ctx2.tuning == ctx.tuning
    # force re-times
    r3 = at.tune(ctx2, lambda: None, key="k1", timer=timer, force=True)
    assert not r3["cached"]


def test_in_memory_cache_is_not_written(tmp_path, monkeypatch):
    monkeypatch.setenv("EMPM_TUNING_CACHE", str(tmp_path / "tuning.json"))
    ctx = FakeCtx()
    cache = {}
    at.tune(ctx, lambda: None, key="k2", cache=cache, timer=lambda run, reps: cost(ctx.tuning))
    assert "k2" in cache and not (tmp_path / "tuning.json").exists()


def test_workload_keys_separate_workloads():
    k = at.workload_key("cpu", [64], 200_000, [7], 1, True)
    assert k == at.workload_key("cpu", [64], 210_000, [7], 1, True)  # same power of two
    others = [at.workload_key("cpu", [32], 200_000, [7], 1, True), at.workload_key("cpu", [64], 50_000, [7], 1, True),
              at.workload_key("cpu", [64], 200_000, [2, 5], 1, True), at.workload_key("cpu", [64], 200_000, [7], 32, True),
              at.workload_key("cpu", [64], 200_000, [7], 1, False)]
    assert len(set(others + [k])) == len(others) + 1


def test_unreadable_cache_is_empty(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text("{not json")
    assert at.load_cache(str(p)) == {}
    assert at.load_cache(str(tmp_path / "missing.json")) == {}
