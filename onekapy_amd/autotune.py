"""Autotuner of the pool tracer's launch shape for a batch of sites.

The knobs are the realization chunks in flight per wave (pool_slots), the
resident waves per CU (pool_waves_per_cu, which also sets each wave's LDS
share and so its tile pool) and the particles per queue item (pool_chunk).
The search is coordinate descent from the library's defaults, one knob at a
time, timing replays of the batch's captured run, repeated until a sweep over
all knobs changes nothing (the knobs interact: fewer slots leave room for
more waves); knob sets that do not fit the LDS are skipped.  Results are cached per batch signature and device in a
JSON file.
"""
import json
import os

from . import _lib

SEARCH = (
    ("pool_chunk", (32, 64, 128, 256)),
    ("pool_waves_per_cu", (4, 6, 8, 12, 16)),
    ("pool_slots", (2, 3, 4, 6)),
)
DEFAULT = {"pool_slots": 2, "pool_waves_per_cu": 16, "pool_chunk": 64}


def signature(batch):
    """The batch's shape as seen by the tuner: sites, realizations, particles,
    largest well count and grid, and the device's compute units."""
    sites = batch.sites
    info = batch.ctx.device_info()
    return "|".join(str(v) for v in (
        len(sites), sum(batch.nreal), max(s.npaths for s in sites),
        max(len(s.wells) for s in sites), max(s.nrows * s.ncols for s in sites),
        info["num_cu"]))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def cached(batch, path):
    """The cached knobs for this batch's signature, or None."""
    return _load(path).get(signature(batch)) if path else None


def tune_batch(batch, repeats=1, cache_path=None, search=SEARCH, log=None, passes=3,
               start=None):
    """Tunes `batch` in place; returns (best knobs, [(knobs, ms or None)]).
    The search starts from `start` (default DEFAULT); at most `passes`
    sweeps; a knob set is timed once."""
    best = dict(DEFAULT if start is None else start)
    trials = []
    seen = {}

    def timed(knobs):
        key = tuple(sorted(knobs.items()))
        if key in seen:
            return seen[key]
        try:
            batch.tune(**knobs)
            ms = min(batch.run() for _ in range(max(1, repeats)))
        except _lib.OkaError:
            ms = None   # does not fit the LDS at this shape
        seen[key] = ms
        trials.append((dict(knobs), ms))
        if log:
            log(f"autotune {knobs}: {ms if ms is None else round(ms, 2)} ms")
        return ms

    best_ms = timed(best)
    for _ in range(max(1, passes)):
        start = dict(best)
        for name, values in search:
            for v in values:
                cand = dict(best, **{name: v})
                ms = timed(cand)
                if ms is not None and (best_ms is None or ms < best_ms):
                    best, best_ms = cand, ms
        if best == start:
            break
    batch.tune(**best)
    if cache_path:
        data = _load(cache_path)
        data[signature(batch)] = dict(best, ms=best_ms)
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    return best, trials
