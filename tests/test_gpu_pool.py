"""The persistent pool schedule in both arithmetics and for both velocity
fields: each trace's arithmetic is lane-local, so the pool must give the
workgroup tracer's counts, per-trace statuses and statistics exactly, and in
FP64-exact mode the oracle's counts bit for bit."""
import numpy as np
import pytest

from onekapy_amd import Realizations, _lib, chronicle_into_field, sites, start_points
from onekapy_amd.sites import fixed_field
from helpers import oracle_chronicle, oracle_fit_draw, small_site

pytestmark = pytest.mark.gpu

STAT_KEYS = ("attempted", "accepted", "visited", "escaped", "candidates")


def _both(s, d, coef, pf, umbra, precision, confined=True, **pool):
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    args = (pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration, s.tol, s.maxstep, umbra)
    wg = chronicle_into_field(*args, precision=precision, confined=confined)
    pl = chronicle_into_field(*args, precision=precision, confined=confined,
                              schedule=_lib.SCHED_POOL, **pool)
    assert np.array_equal(pl.counts, wg.counts)
    assert np.array_equal(pl.trace_status, wg.trace_status)
    for k in STAT_KEYS:
        assert getattr(pl.stats, k) == getattr(wg.stats, k), k
    assert list(pl.stats.status_count) == list(wg.stats.status_count)
    return pl, wg


POOL_CASES = [
    ("default", {}),
    ("whole_realizations", dict(pool_chunk=1000)),
    ("odd_chunks_one_slot", dict(pool_chunk=7, pool_slots=1)),
    ("tiny_pool", dict(pool_slots=4, pool_waves_per_cu=8)),
]


@pytest.mark.parametrize("name,pool", POOL_CASES, ids=[c[0] for c in POOL_CASES])
def test_fp64_pool_matches_workgroup_and_oracle_config1(gpu_ctx, name, pool):
    """FP64-exact on the pool (configuration 1, 24 realizations x 1000
    particles) equals the FP64-exact workgroup tracer and the oracle."""
    print(f"fp64 pool case {name}", flush=True)
    s = sites.config1(24)
    d = sites.draw(s, seed=13)
    _, _, coef = oracle_fit_draw(s, d)
    pf = s.fixed_field()
    pl, _ = _both(s, d, coef, pf, s.umbra, _lib.FP64_EXACT, **pool)
    if name == "default":
        ocounts, ost = oracle_chronicle(s, d, coef, pf)
        assert np.array_equal(pl.counts, ocounts)
        assert (pl.stats.attempted, pl.stats.accepted, pl.stats.visited) == (ost[0], ost[1],
                                                                              ost[7])


def test_fp64_pool_many_tiles_and_uneven_lifetimes(gpu_ctx):
    """A wide umbra on a fine 1200 x 1100 lattice (tiles beyond a slot's map,
    in HBM) with uneven lifetimes, FP64-exact: pool = workgroup = oracle."""
    s, d = small_site(nreal=8, npaths=40)
    _, _, coef = oracle_fit_draw(s, d)
    pf = fixed_field(s.xt, s.yt, 5.0, 1100, 1200)
    for chunk in (40, 5):
        print(f"chunk {chunk}", flush=True)
        pl, _ = _both(s, d, coef, pf, 300.0, _lib.FP64_EXACT, pool_slots=3, pool_chunk=chunk)
    s.umbra = 300.0
    ocounts, _ = oracle_chronicle(s, d, coef, pf)
    assert np.array_equal(pl.counts, ocounts)


@pytest.mark.parametrize("precision", [_lib.FP64_EXACT, _lib.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("shift", [0.0, 0.7])
def test_unconfined_pool_matches_workgroup(gpu_ctx, precision, shift):
    """The unconfined velocity on the pool (configuration 0, F lowered by
    shift * K b^2 / 2: at 0.7 a fifth of the realizations are dry at their
    start points): the workgroup tracer's counts, statuses and statistics,
    and in FP64 the oracle's counts."""
    print(f"unconfined pool precision {precision} shift {shift}", flush=True)
    s = sites.config0()
    d = sites.draw(s, seed=11)
    _, _, coef = oracle_fit_draw(s, d)
    coef[:, 5] -= shift * 0.5 * d["K"] * d["b"] ** 2
    pf = s.fixed_field()
    pl, _ = _both(s, d, coef, pf, s.umbra, precision, confined=False, pool_chunk=32)
    if shift > 0:
        assert 0 < pl.stats.status_count[_lib.TRACE_TRUNCATED] < s.nrealizations * s.npaths
    if precision == _lib.FP64_EXACT:
        ocounts, ost = oracle_chronicle(s, d, coef, pf, confined=False)
        assert np.array_equal(pl.counts, ocounts)
        assert list(pl.stats.status_count) == list(ost[2:6])
