"""Autotuner logic on the host: a fake context whose workload time is a
known function of the tuning, so the search must find its minimum."""
import pytest

from mtimager_amd import autotune


class _FakeCtx:
    def __init__(self, cost):
        self.cost = cost
        self.t = dict(autotune.DEFAULTS)
        self.sets = []

    def set_tuning(self, **kw):
        self.t.update(kw)
        self.sets.append(dict(self.t))


class _Clock:
    """Advances by the current tuning's cost on every run."""

    def __init__(self, ctx):
        self.ctx, self.now, self.runs = ctx, 0.0, 0

    def __call__(self):
        return self.now

    def run(self):
        self.runs += 1
        self.now += self.ctx.cost(self.ctx.t)


SPACE = {"unit_pairs": [512, 1024, 2048, 4096, 8192], "whole_columns_ny": [0, 256]}


def _cost(t):
    # separable bowl: best unit_pairs 4096, best whole_columns_ny 0
    u = {512: 5.0, 1024: 3.0, 2048: 2.0, 4096: 1.0, 8192: 1.5}[t["unit_pairs"]]
    w = {0: 0.1, 256: 0.4}[t["whole_columns_ny"]]
    return u + w


@pytest.mark.parametrize("strategy", ["grid", "coordinate"])
def test_tuner_finds_the_fastest_and_leaves_it_set(strategy):
    ctx = _FakeCtx(_cost)
    clk = _Clock(ctx)
    best, timings = autotune.tune(ctx, clk.run, space=SPACE, repeats=2, strategy=strategy, sync=lambda: None,
                                  clock=clk)
    assert best == {"unit_pairs": 4096, "whole_columns_ny": 0}
    assert all(ctx.t[k] == v for k, v in best.items())
    assert all(abs(t - _cost(c)) < 1e-12 for c, t in timings)
    n = len(timings)
    assert n == (10 if strategy == "grid" else 6)
    assert clk.runs == n * 3  # one warm-up and two timed runs per candidate


def test_coordinate_search_keeps_earlier_choices():
    # the second knob's sweep runs at the first knob's best value
    ctx = _FakeCtx(_cost)
    clk = _Clock(ctx)
    _, timings = autotune.tune(ctx, clk.run, space=SPACE, repeats=1, sync=lambda: None, clock=clk)
    second = [c for c, _ in timings if c["whole_columns_ny"] == 0]
    assert second and all(c["unit_pairs"] == 4096 for c in second)


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tuning.json")
    c = autotune.TuningCache(path)
    key = autotune.TuningCache.key("MI355X", "snapshot-mix")
    ctx = _FakeCtx(_cost)
    clk = _Clock(ctx)
    t, timings = autotune.tune_cached(ctx, clk.run, c, key, space=SPACE, repeats=1, sync=lambda: None, clock=clk)
    assert timings is not None and t == {"unit_pairs": 4096, "whole_columns_ny": 0}
    # a new process reads the file and sets the tuning without timing anything
    ctx2 = _FakeCtx(_cost)
    clk2 = _Clock(ctx2)
    t2, timings2 = autotune.tune_cached(ctx2, clk2.run, autotune.TuningCache(path), key, sync=lambda: None,
                                        clock=clk2)
    assert timings2 is None and t2 == t and all(ctx2.t[k] == v for k, v in t.items()) and clk2.runs == 0


def test_default_space_covers_every_knob():
    assert set(autotune.DEFAULT_SPACE) == set(autotune.DEFAULTS) == {"unit_pairs", "whole_columns_ny", "xcd_chunk",
                                                                     "region_width"}
    assert all(autotune.DEFAULTS[k] in v for k, v in autotune.DEFAULT_SPACE.items())


def test_unknown_strategy():
    with pytest.raises(ValueError):
        autotune.tune(_FakeCtx(_cost), lambda: None, strategy="random", sync=lambda: None)
