"""The launch-shape autotuner's search logic, on a stand-in batch (no GPU)."""
import json

from onekapy_amd import _lib
from onekapy_amd.autotune import DEFAULT, cached, tune_batch


class _Site:
    npaths, wells, nrows, ncols = 256, [0] * 7, 100, 100


class _Ctx:
    def device_info(self):
        return {"num_cu": 256}


class FakeBatch:
    """Replay time is a function of the knobs.  The fastest shape needs two
    knobs changed together: more waves per CU only pay once the slots are
    cut to 2 (as with the LDS share of a wave), and 16 waves with 4 slots do
    not fit."""

    def __init__(self):
        self.sites, self.nreal, self.ctx = [_Site()], [100], _Ctx()
        self.knobs = {"pool_slots": 4, "pool_waves_per_cu": 8, "pool_chunk": 64}
        self.runs = 0

    def tune(self, **knobs):
        k = dict(self.knobs, **knobs)
        if k["pool_slots"] * k["pool_waves_per_cu"] > 48:
            raise _lib.OkaError(-2, "does not fit the LDS")
        self.knobs = k

    def run(self):
        self.runs += 1
        k = self.knobs
        t = 100.0 + abs(k["pool_chunk"] - 64) / 8
        t += 20.0 * abs(k["pool_slots"] - 2) if k["pool_waves_per_cu"] >= 12 else 0.0
        t -= 2.0 * k["pool_waves_per_cu"] if k["pool_slots"] <= 2 else 0.0
        t += 1.0 * (k["pool_slots"] < 4)
        return t


def test_coordinate_descent_needs_a_second_sweep(tmp_path):
    start = {"pool_slots": 4, "pool_waves_per_cu": 8, "pool_chunk": 64}
    b = FakeBatch()
    one, _ = tune_batch(b, passes=1, start=start)
    b2 = FakeBatch()
    path = str(tmp_path / "tune.json")
    best, trials = tune_batch(b2, cache_path=path, start=start)
    # one sweep stops at 2 slots x 8 waves; the second finds 2 slots x 16 waves
    assert one["pool_waves_per_cu"] == 8 and one["pool_slots"] == 2
    assert best == {"pool_slots": 2, "pool_waves_per_cu": 16, "pool_chunk": 64}
    assert b2.knobs == best                      # the batch is left tuned
    keys = [tuple(sorted(k.items())) for k, _ in trials]
    assert len(keys) == len(set(keys))           # every shape is timed once
    assert any(ms is None for _, ms in trials)   # shapes that do not fit are skipped
    assert cached(b2, path)["pool_waves_per_cu"] == 16
    assert json.load(open(path))


def test_default_start_is_a_fitting_shape():
    b = FakeBatch()
    best, trials = tune_batch(b)
    assert trials[0][0] == DEFAULT and trials[0][1] is not None
