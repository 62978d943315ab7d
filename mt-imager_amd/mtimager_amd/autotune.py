"""Autotuner for a context's launch-shape choices (mti_tuning): the gridder's
region shape (each wave of a tile's workgroup owns 4 x 16, 8 x 8 or 16 x 4
cells -- a kernel variant each: wider regions read longer runs of a sample
row, narrower ones waste fewer lanes on small support boxes), its work-unit
size (how many (record, tile) pairs one workgroup takes, which sets how a
crowded tile splits over workgroups and how many tiles need atomic flushes),
how work units are dealt to the 8 XCDs (runs of neighbouring tiles per XCD,
for L2 reuse of kernel samples) and the transform's column strategy where two
apply.

Each candidate is set on the context, run once untimed, then timed `repeats`
times on the caller's representative workload; the fastest stays set.  The
search is either the full grid of the space or coordinate descent (one knob
at a time, the others at their best so far).  Results can be kept in a JSON
file keyed by device and workload, so a service tunes once per workload class.

The reference has no tuner: its gridder blocks are fixed at <= 440 / <= 110
records (op/common.hcu:26-27).
"""
import itertools
import json
import os
import time

# the knobs of mti_tuning and the values searched by default
DEFAULT_SPACE = {
    "region_width": [8, 16, 4],
    "xcd_chunk": [0, 8, 32, 128],
    "unit_pairs": [512, 1024, 2048, 4096, 8192],
    "whole_columns_ny": [0, 256],
}
DEFAULTS = {"unit_pairs": 2048, "whole_columns_ny": 256, "xcd_chunk": 0, "region_width": 8}


def grid_candidates(space):
    keys = list(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, vals))


def _time(ctx, run, sync, clock, cand, repeats):
    ctx.set_tuning(**cand)
    run()
    sync()
    best = float("inf")
    for _ in range(repeats):
        t0 = clock()
        run()
        sync()
        best = min(best, clock() - t0)
    return best


def tune(ctx, run, space=None, repeats=3, strategy="coordinate", sync=None, clock=time.perf_counter):
    """Times `run()` (the workload, enqueued on ctx) under each candidate
    tuning and leaves the fastest set on ctx.  sync() waits for the device
    (torch.cuda.synchronize by default).  Returns (best, timings) with
    timings a list of (candidate, seconds)."""
    space = dict(DEFAULT_SPACE if space is None else space)
    if sync is None:
        import torch
        sync = torch.cuda.synchronize
    timings = []
    if strategy == "grid":
        for cand in grid_candidates(space):
            timings.append((cand, _time(ctx, run, sync, clock, cand, repeats)))
        best = min(timings, key=lambda t: t[1])[0]
    elif strategy == "coordinate":
        best = {k: (DEFAULTS[k] if DEFAULTS.get(k) in vals else vals[0]) for k, vals in space.items()}
        seen = {}
        for k, vals in space.items():
            for v in vals:
                cand = dict(best, **{k: v})
                key = tuple(sorted(cand.items()))
                if key not in seen:
                    seen[key] = _time(ctx, run, sync, clock, cand, repeats)
                    timings.append((cand, seen[key]))
            best = min((c for c, _ in timings if all(c[j] == best[j] for j in best if j != k)),
                       key=lambda c: seen[tuple(sorted(c.items()))])
    else:
        raise ValueError(f"unknown strategy {strategy}")
    ctx.set_tuning(**best)
    return best, timings


class TuningCache:
    """Tuning results on disk, {key: {"unit_pairs": .., "whole_columns_ny": ..}}."""

    def __init__(self, path):
        self.path = path
        self.table = {}
        if os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    @staticmethod
    def key(device_name, workload):
        return f"{device_name}|{workload}"

    def get(self, key):
        t = self.table.get(key)
        return dict(t) if t is not None else None

    def put(self, key, tuning):
        self.table[key] = {k: int(v) for k, v in tuning.items()}

    def save(self):
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def tune_cached(ctx, run, cache, key, **kw):
    """The cached tuning for `key` if there is one (set on ctx), else tune()
    and record the result.  Returns (tuning, timings or None)."""
    t = cache.get(key)
    if t is not None:
        ctx.set_tuning(**t)
        return t, None
    best, timings = tune(ctx, run, **kw)
    cache.put(key, best)
    cache.save()
    return best, timings
