"""Autotuner for the launch parameters of batched and single-bunch pushes.

Coordinate search over block size, steps per launch and cell-sort cadence, each
candidate timed on a short sample of the real workload (the same bunches, a
few hundred steps, state re-uploaded for every candidate).  Results are cached
per process, keyed by a workload signature (precision, particle count, bunch
count, map shape), so a serving process tunes once; set PYPAT_AUTOTUNE_CACHE to
a JSON file path to keep them across processes.  The launch parameters change
the schedule only: every candidate computes the same results bit for bit.

    cfg = autotune_batch(maps, bunches, "fp32_mixed")
    BatchedTracker(maps, "fp32_mixed", **cfg)
"""
import json
import os
import time

from .batch import BatchedTracker

BLOCKS = (128, 256)
STEPS_PER_LAUNCH = (250, 500, 1000)
SORT_EVERY = (0, 250, 500, 1000, 2000)
DEFAULT = {"block_size": 256, "steps_per_launch": 500, "sort_every": 500}


_MEMORY = {}


def cache_path():
    """The JSON file of tuned configurations (PYPAT_AUTOTUNE_CACHE), or None: this process only."""
    return os.environ.get("PYPAT_AUTOTUNE_CACHE") or None


def _load_cache():
    path = cache_path()
    if path is None:
        return dict(_MEMORY)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache):
    _MEMORY.update(cache)
    path = cache_path()
    if path is None:
        return
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass


def signature(maps, bunches, precision):
    n = sum(b.n for b in bunches)
    shape = "x".join(str(s) for s in maps[0].shape) if hasattr(maps[0], "shape") else "?"
    return f"{precision}|n={n}|bunches={len(bunches)}|map={shape}"


def _time_config(maps, bunches, precision, cfg, sample_steps):
    tr = BatchedTracker(maps, precision, **cfg).load(bunches)
    tr.run(nsteps=min(50, sample_steps))          # warm the code and caches
    tr.load(bunches)
    t0 = time.perf_counter()
    tr.run(nsteps=sample_steps)
    return time.perf_counter() - t0


def autotune_batch(maps, bunches, precision="fp32_mixed", sample_steps=1000, use_cache=True, verbose=False):
    """Best {block_size, steps_per_launch, sort_every} for these bunches (coordinate search)."""
    sig = signature(maps, bunches, precision)
    cache = _load_cache() if use_cache else {}
    if sig in cache:
        return dict(cache[sig]["config"])
    best = dict(DEFAULT)
    table = []

    def trial(cfg):
        t = _time_config(maps, bunches, precision, cfg, sample_steps)
        table.append((dict(cfg), t))
        if verbose:
            print(f"autotune {cfg} -> {t * 1e3:.1f} ms", flush=True)
        return t

    best_t = trial(best)
    # a sample must show what a candidate costs over a long run: a sort cadence is tried only where the
    # sample holds at least two of its sorts (never sorting only where it spans 4000 steps), a launch
    # length only where the sample holds one launch of it; otherwise a cadence that skips its sorts in a
    # short sample looks cheapest and then slows the long run
    sorts = tuple(v for v in SORT_EVERY if (v == 0 and sample_steps >= 4000) or (v > 0 and 2 * v <= sample_steps))
    lengths = tuple(v for v in STEPS_PER_LAUNCH if v <= sample_steps)
    for key, values in (("sort_every", sorts), ("block_size", BLOCKS), ("steps_per_launch", lengths)):
        for val in values:
            if val == best[key]:
                continue
            cfg = dict(best)
            cfg[key] = val
            t = trial(cfg)
            if t < best_t:
                best, best_t = cfg, t
    if use_cache:
        cache[sig] = {"config": best, "seconds": best_t, "sample_steps": sample_steps,
                      "trials": [[c, t] for c, t in table]}
        _store_cache(cache)
    return best
