"""Many sites in one graph-captured run (oka_batch_*) against per-site runs,
and the autotuner of its launch shape."""
import numpy as np
import pytest

from onekapy_amd import _lib, fit_and_draw, sites
from onekapy_amd.autotune import cached, tune_batch
from onekapy_amd.batch import SiteBatch
from onekapy_amd.surrogate import trace_counts

pytestmark = pytest.mark.gpu


def _small_batch():
    ss = sites.config2(nsites=6, scale=0.004, seed=7)
    draws = [sites.draw(s, seed=11 + i) for i, s in enumerate(ss)]
    return ss, draws


def test_batch_matches_per_site_runs(gpu_ctx):
    """Every site of a batch with far-well expansions off (far_order -1) equals
    the workgroup tracer run on that site alone."""
    ss, draws = _small_batch()
    b = SiteBatch(ss, draws, far_order=-1)
    b.run()
    for i, (s, d) in enumerate(zip(ss, draws)):
        print(f"site {i}: {s.name}", flush=True)
        counts, coef, st = b.results(i)
        _, _, ref_coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"],
                                      d["b"], d["q"], d["z"])
        assert np.array_equal(coef, ref_coef)
        # the same tracer, one site at a time, on the site's fixed lattice (the
        # batch drops what falls outside it, as a pre-expanded field does)
        ref_counts, ref_st = trace_counts(s, d, ref_coef, _lib.FP32, far_order=-1)
        assert np.array_equal(counts, ref_counts)
        assert (st.attempted, st.accepted, st.visited, st.escaped) == (
            ref_st.attempted, ref_st.accepted, ref_st.visited, ref_st.escaped)
        assert counts.max() <= len(d["K"])
    fields = b.fields()
    assert fields[0].total_weight == len(draws[0]["K"])
    b.close()


def test_batch_far_expansions_match_single_site_pool(gpu_ctx):
    """With the library's choice of far-well expansions (sites of 64 wells or
    more), each site of the batch equals the pool tracer run on that site alone
    with the same choice: the batch stages the same box plan, and each slot's
    coefficients depend only on its realization."""
    ss, draws = _small_batch()
    assert max(len(s.wells) for s in ss) >= 64
    b = SiteBatch(ss, draws)
    b.run()
    for i, (s, d) in enumerate(zip(ss, draws)):
        print(f"site {i}: {len(s.wells)} wells", flush=True)
        counts, coef, st = b.results(i)
        ref_counts, ref_st = trace_counts(s, d, coef, _lib.FP32, _lib.SCHED_POOL)
        assert np.array_equal(counts, ref_counts), i
        assert (st.attempted, st.accepted, st.visited) == (
            ref_st.attempted, ref_st.accepted, ref_st.visited), i
    b.close()


def test_config2_full_size_batch_matches_per_site_runs(gpu_ctx):
    """The 64-site uneven workload at its full size (about 1.5 * 10^5
    realizations x 256 particles) in one graph replay: every site's counts and
    step statistics equal a run of that site alone, and the counts stay within
    the site's realization count."""
    ss = sites.config2()
    draws = [sites.draw(s, seed=100 + i) for i, s in enumerate(ss)]
    b = SiteBatch(ss, draws)
    b.run()
    total = 0
    for i, (s, d) in enumerate(zip(ss, draws)):
        print(f"site {i}: {s.name}", flush=True)
        counts, coef, st = b.results(i)
        ref_counts, ref_st = trace_counts(s, d, coef, _lib.FP32, _lib.SCHED_POOL)
        assert np.array_equal(counts, ref_counts), i
        assert (st.attempted, st.accepted, st.visited) == (
            ref_st.attempted, ref_st.accepted, ref_st.visited), i
        assert counts.max() <= len(d["K"]) and counts.sum() == st.visited
        total += len(d["K"])
    assert total > 100000
    b.close()


def test_batch_replays_and_retunes_identically(gpu_ctx):
    """Every replay of the captured graph (fills, fits, pool launch) equals the
    eager per-site runs, with eager fits and traces on the device between the
    replays."""
    ss, draws = _small_batch()
    eager = []
    for s, d in zip(ss, draws):
        _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"],
                                  d["b"], d["q"], d["z"])
        eager.append(trace_counts(s, d, coef, _lib.FP32, _lib.SCHED_POOL)[0])
    b = SiteBatch(ss, draws)
    for rep in range(4):                      # counts restart from zero on every replay
        print(f"replay {rep}", flush=True)
        b.run()
        for i in range(len(ss)):
            assert np.array_equal(b.results(i)[0], eager[i]), (rep, i)
        trace_counts(ss[rep], draws[rep], fit_and_draw(
            ss[rep].wells_xy(), ss[rep].xt, ss[rep].yt, ss[rep].base, ss[rep].observations,
            draws[rep]["K"], draws[rep]["b"], draws[rep]["q"], draws[rep]["z"])[2], _lib.FP32)
    first = eager
    for knobs in (dict(pool_chunk=16, pool_slots=2), dict(pool_chunk=256, pool_waves_per_cu=4)):
        print(f"knobs {knobs}", flush=True)
        b.tune(**knobs)
        b.run()
        for i in range(len(ss)):
            assert np.array_equal(b.results(i)[0], first[i])
    b.close()


def test_autotuner_picks_a_fitting_shape_and_caches_it(gpu_ctx, tmp_path):
    ss, draws = _small_batch()
    b = SiteBatch(ss, draws)
    path = str(tmp_path / "tune.json")
    search = (("pool_chunk", (64, 256)), ("pool_slots", (2, 16)))
    best, trials = tune_batch(b, cache_path=path, search=search)
    times = [ms for _, ms in trials if ms is not None]
    assert times and min(times) > 0
    assert cached(b, path)["pool_chunk"] == best["pool_chunk"]
    # the tuned batch still gives the same counts
    b.run()
    ref = SiteBatch(ss, draws)
    ref.run()
    for i in range(len(ss)):
        assert np.array_equal(b.results(i)[0], ref.results(i)[0])
    b.close()
    ref.close()


# configuration-2 sites with few wells and short traces (so the oracle runs
# them in seconds), and a mix with up to 135 wells and 36-year traces
ORACLE_SITES = (44, 33)
MIXED_SITES = (2, 7, 12, 33, 44, 53)


def _cfg2_subset(idx, nreal):
    ss = sites.config2()
    out, draws = [], []
    for i in idx:
        s = ss[i]
        d = sites.draw(s, seed=100 + i)
        n = min(nreal, len(d["K"]))
        out.append(s)
        draws.append({k: v[:n] for k, v in d.items()})
        s.nrealizations = n
    return out, draws


def test_config2_batch_within_bound_of_fp64_exact(gpu_ctx):
    """Six configuration-2 sites (1 to 135 wells, 4.9 to 35.7 years, 144 to
    670^2 nodes), first 300 realizations each, in one batch replay (FP32 pool
    tracer): every site's capture probabilities are within 0.01 of the FP64-exact
    tracer's on the same realizations, at every node."""
    ss, draws = _cfg2_subset(MIXED_SITES, 300)
    b = SiteBatch(ss, draws)
    b.run()
    for i, (s, d) in enumerate(zip(ss, draws)):
        counts, coef, st = b.results(i)
        ref, ref_st = trace_counts(s, d, coef, _lib.FP64_EXACT)
        R = len(d["K"])
        err = np.abs(counts - ref) / R
        print(f"{s.name}: R {R}, max |p32 - p64| {err.max():.4f}, L1 {err.sum():.3f}, "
              f"steps {st.attempted} vs {ref_st.attempted}", flush=True)
        assert ref.sum() > 0
        assert err.max() <= 0.01
        assert err.sum() <= 0.01 * (ref > 0).sum()
    b.close()


def test_config2_sites_fp64_exact_match_oracle(gpu_ctx):
    """Two configuration-2 sites (6 and 8 wells), first 8 realizations: the
    batch's fits and MVN draws equal the oracle's, and the FP64-exact tracer on
    those coefficients equals the oracle bit for bit (counts, steps, statuses)."""
    from helpers import oracle_chronicle, oracle_fit_draw
    ss, draws = _cfg2_subset(ORACLE_SITES, 8)
    b = SiteBatch(ss, draws)
    b.run()
    for i, (s, d) in enumerate(zip(ss, draws)):
        print(f"{s.name}: oracle", flush=True)
        _, coef, _ = b.results(i)
        _, _, ocoef = oracle_fit_draw(s, d)
        assert np.array_equal(coef, ocoef)
        counts, st = trace_counts(s, d, coef, _lib.FP64_EXACT)
        ocounts, ost = oracle_chronicle(s, d, coef, s.fixed_field())
        assert np.array_equal(counts, ocounts)
        assert (st.attempted, st.accepted, st.visited, st.escaped) == (ost[0], ost[1], ost[7],
                                                                       ost[6])
        assert list(st.status_count) == list(ost[2:6])
        assert counts.sum() > 0
    b.close()


def test_fp64_batch_matches_oracle(gpu_ctx):
    """A batch in FP64-exact (the pool tracer in the workgroup tracer's
    arithmetic): every site's counts and statistics equal the oracle's on the
    batch's own fits, which equal the oracle's fits."""
    from helpers import oracle_chronicle, oracle_fit_draw
    ss, draws = _cfg2_subset((44, 33, 7), 16)
    b = SiteBatch(ss, draws, precision=_lib.FP64_EXACT)
    b.run()
    for i, (s, d) in enumerate(zip(ss, draws)):
        print(f"{s.name}: fp64 batch vs oracle", flush=True)
        counts, coef, st = b.results(i)
        _, _, ocoef = oracle_fit_draw(s, d)
        assert np.array_equal(coef, ocoef)
        ocounts, ost = oracle_chronicle(s, d, coef, s.fixed_field())
        assert np.array_equal(counts, ocounts)
        assert (st.attempted, st.accepted, st.visited, st.escaped) == (ost[0], ost[1], ost[7],
                                                                       ost[6])
        assert list(st.status_count) == list(ost[2:6])
    b.close()


@pytest.mark.parametrize("precision", [_lib.FP32, _lib.FP64_EXACT], ids=["fp32", "fp64"])
def test_batch_mixes_confined_and_unconfined_sites(gpu_ctx, precision):
    """Sites tracing the unconfined velocity beside confined ones in one batch
    (the pool's per-slot choice of field): each site's counts and statistics
    equal that site traced alone with its own field."""
    ss, draws = _cfg2_subset((44, 33, 7, 12), 24)
    ss[1].confined = False
    ss[3].confined = False
    b = SiteBatch(ss, draws, precision=precision)
    b.run()
    for i, (s, d) in enumerate(zip(ss, draws)):
        print(f"{s.name}: confined {s.confined}", flush=True)
        counts, coef, st = b.results(i)
        ref, ref_st = trace_counts(s, d, coef, precision, _lib.SCHED_POOL)
        assert np.array_equal(counts, ref)
        assert (st.attempted, st.accepted, st.visited) == (ref_st.attempted, ref_st.accepted,
                                                            ref_st.visited)
        assert list(st.status_count) == list(ref_st.status_count)
    b.close()
