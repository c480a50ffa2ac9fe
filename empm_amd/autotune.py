"""Autotuner of the binned kernels' launch parameters (empm_tuning).

The block-owned scatters (P2G and the G2P adjoint's dL/dv_out scatter) and
the grid node passes have launch parameters whose best values depend on the
workload: how many waves share one block's LDS tile (the scatter workgroup
size), how many scatter workgroups run per CU, which blocks are scattered
first (blocks of more than `heavy_chunks` 64-particle chunks), when blocks
are handed out by a work counter instead of in fixed strides, and how wide
a node pass is.  Results do not depend on them beyond the order of float
atomic sums.

The tuner times real substeps of the caller's own workload (a forward
rollout, and with `backward` its reverse sweep) under each candidate and
walks the parameters by coordinate descent: every value of one parameter
with the others fixed, keeping the fastest, parameter after parameter,
until a pass changes nothing.  A candidate replaces the current one only
when it is faster by more than `rel_tol` (timing noise).  Results are cached
per workload signature (device, grids, particle count to the nearest power
of two, materials, simulations) in a JSON file, so a later run of the same
workload sets its tuning without timing anything.

No counterpart in the reference (Warp picks its own launch shapes); the
workload signature and search are this library's own.
"""
from __future__ import annotations

import json
import math
import os
import statistics

import torch

DEFAULT = {"scatter_wg": 256, "scatter_per_cu": 0, "heavy_chunks": 8, "dynamic_rounds": 4, "node_grid_cap": 4096}

# the values searched per parameter (the defaults included)
SPACE = {
    "scatter_wg": (128, 256, 512),
    "scatter_per_cu": (0, 1, 2),
    "heavy_chunks": (0, 4, 8, 16),
    "dynamic_rounds": (1, 2, 4, 8),
    "node_grid_cap": (1024, 2048, 4096, 8192, 16384),
}


def coordinate_search(measure, start=None, space=None, passes=3, rel_tol=0.01):
    """Coordinate descent over `space` from `start`; measure(cfg) -> seconds.
    Returns (best config, its time, [(config, time), ...] of every timing)."""
    space = SPACE if space is None else space
    best = dict(DEFAULT if start is None else start)
    best_t = measure(dict(best))
    trace = [(dict(best), best_t)]
    for _ in range(passes):
        changed = False
        for key, values in space.items():
            for v in values:
                if v == best[key]:
                    continue
                cand = dict(best)
                cand[key] = v
                t = measure(dict(cand))
                trace.append((dict(cand), t))
                if t < best_t * (1.0 - rel_tol):
                    best, best_t, changed = cand, t, True
        if not changed:
            break
    return best, best_t, trace


def workload_key(device, grids, n_particles, materials, n_sims=1, backward=True):
    """Signature of a workload for the tuning cache."""
    dev = torch.device(device)
    name = torch.cuda.get_device_name(dev) if dev.type == "cuda" else dev.type
    grids = sorted(set(int(g) for g in grids))
    nb = int(round(math.log2(max(int(n_particles), 1))))
    mats = sorted(set(int(m) for m in materials))
    return f"{name}|grids={grids}|N~2^{nb}|mat={mats}|sims={int(n_sims)}|{'fwd+bwd' if backward else 'fwd'}"


def cache_path():
    return os.environ.get("EMPM_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "empm_amd", "tuning.json"))


def load_cache(path=None):
    try:
        with open(path or cache_path()) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=None):
    path = path or cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


def time_call(fn, reps=3, warmup=1):
    """Median seconds of fn() over reps runs, timed with events on the
    current stream after warmup runs."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b) / 1000.0)
    return statistics.median(ts)


def tune(ctx, run, key=None, cache=None, force=False, reps=3, space=None, passes=3, rel_tol=0.01, timer=None):
    """Tune context `ctx` for the workload run() executes (a rollout on ctx,
    forward and optionally backward).  With `key` the result is looked up in
    and stored to the cache (a dict, saved by the caller, or None for the
    default file).  timer(run, reps) -> seconds (default: time_call).
    Leaves the best tuning set on ctx and returns {"tuning", "seconds",
    "default_seconds", "timings", "cached"}."""
    timer = time_call if timer is None else timer
    own_cache = cache is None
    if own_cache:
        cache = load_cache()
    if key is not None and not force and key in cache:
        ent = cache[key]
        ctx.set_tuning(ent["tuning"])
        return dict(ent, cached=True)

    def measure(cfg):
        ctx.set_tuning(cfg)
        return timer(run, reps)

    best, best_t, trace = coordinate_search(measure, space=space, passes=passes, rel_tol=rel_tol)
    ctx.set_tuning(best)
    default_t = next(t for c, t in trace if c == DEFAULT) if any(c == DEFAULT for c, _ in trace) else None
    res = {"tuning": best, "seconds": best_t, "default_seconds": default_t, "timings": len(trace)}
    if key is not None:
        cache[key] = res
        if own_cache:
            save_cache(cache)
    return dict(res, cached=False)
