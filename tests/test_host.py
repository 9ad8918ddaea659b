"""CPU tests of the product's host side: the C ABI export table, the host build
of the exact scalar math, the expansion-history scan, the ProbabilityField
container, distributions, sites and the loud failure without the library.
No GPU is used here (no compute call reaches a kernel)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import oracle as O
from oracle import pyref
from onekapy_amd import (DistributionError, Lognormal, ProbabilityField, Signed,
                         compute_variate_mean, generate_random_variate, isdistribution, sites,
                         start_points, _lib)
from onekapy_amd.probabilityfield import RangeError
from helpers import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    with open(os.path.join(ROOT, "include", "oka.h")) as f:
        text = f.read()
    declared = set(re.findall(
        r"^\s*(?:int|int64_t|void|double|const char \*|void \*)\s*\**\s*(oka_\w+)\s*\(", text,
        flags=re.M))
    assert declared == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(L, name), name
    assert L.oka_abi_version() == 5


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", os.path.join(ROOT, "does_not_exist.so"))
    monkeypatch.setattr(_lib, "_contexts", {})
    from onekapy_amd import Model
    m = Model(0.0, 1.0, 0.25, 10.0, [(0.0, 0.0, 0.25, 100.0)])
    with pytest.raises(_lib.OkaLibraryError):
        m.compute_potential(1.0, 2.0)


def test_host_exact_math_matches_quad_precision():
    L = _lib.load()
    rng = np.random.default_rng(17)
    xs = np.concatenate([10 ** rng.uniform(-6, 18, 15000), rng.uniform(0.5, 2.0, 5000),
                         2.0 ** rng.integers(-1070, 1023, 300).astype(float),
                         [5e-324, 1e-310, 1.0, 2.0, 32.0, 1.7976931348623157e308]])
    for x in xs:
        x = float(x)
        assert L.oka_pow_fifth(x) == O.pow02(x, O.MATH_CR), x.hex()
        assert L.oka_log(x) == O.log(x, O.MATH_CR), x.hex()
    assert L.oka_pow_fifth(0.0) == 0.0 and math.isnan(L.oka_pow_fifth(-1.0))
    assert L.oka_log(0.0) == -math.inf and math.isnan(L.oka_log(-1.0))


def test_host_log_fast_path_matches_quad_precision():
    """oka_log is the table-driven log with its rounding test (log_rn): against
    the oracle's binary128 logarithm on the table's hard cases -- x within
    2^-6 and 2^-30 of 1 from both sides (no cancellation), every table row's
    boundaries, the upper-half rows' shift, and squared distances of the
    tracers' range -- plus powers of two and subnormals."""
    L = _lib.load()
    rng = np.random.default_rng(29)
    rows = 1.0 + np.arange(129) / 128.0
    xs = np.concatenate([
        1.0 + rng.uniform(-2 ** -7, 2 ** -7, 4000), 1.0 + rng.uniform(-2 ** -30, 2 ** -30, 2000),
        np.nextafter(rows, 0.0), rows[:-1], np.nextafter(rows[:-1], 3.0),
        np.ldexp(rng.uniform(1.0, 2.0, 6000), rng.integers(-300, 300, 6000)),
        rng.uniform(1e-4, 1e8, 6000), 2.0 ** np.arange(-1074, 1024, 7).astype(float),
        [5e-324, 2.2250738585072014e-308, 1.0, 1.7976931348623157e308]])
    for x in xs:
        x = float(x)
        assert L.oka_log(x) == O.log(x, O.MATH_CR), x.hex()
    assert L.oka_log(math.inf) == math.inf


def _oracle_sequence(init, bboxes):
    """Sequential expands (probabilityfield.py:175-261) on the oracle field,
    returning the geometry after each one."""
    f = O.OracleField(init["dx"], init["dy"], *init.get("origin", (float("nan"),) * 2))
    geo = []
    for b in bboxes:
        f.expand(*b)
        geo.append((f.s.xmin, f.s.ymin, f.s.xmax, f.s.ymax, *f.shape))
    return geo


@pytest.mark.parametrize("case", range(12))
def test_grid_history_matches_sequential_expand(case):
    rng = np.random.default_rng(case)
    dx = float(rng.choice([1.0, 0.7, 10.0, 3.3]))
    origin = (float(rng.uniform(-1e4, 1e4)), float(rng.uniform(-1e4, 1e4)))
    empty = case % 3 == 0
    n = 200
    c = np.array(origin) + rng.normal(0, 30 * dx, (n, 2))
    half = np.abs(rng.normal(0, 10 * dx, (n, 2)))
    bbox = np.stack([c[:, 0] - half[:, 0], c[:, 0] + half[:, 0],
                     c[:, 1] - half[:, 1], c[:, 1] + half[:, 1]], axis=1)
    init = {"dx": dx, "dy": dx}
    if empty:
        g0 = _lib.Grid_()
        g0.dx, g0.dy = dx, dx
    else:
        init["origin"] = origin
        g0 = _lib.make_grid(origin[0] - dx, origin[1] - dx, origin[0] + dx, origin[1] + dx, dx, dx,
                            3, 3)
    geo = _oracle_sequence(init, bbox)
    final, epochs = _lib.grid_history(g0, bbox)
    assert (final.xmin, final.ymin, final.xmax, final.ymax, final.nrows, final.ncols) == geo[-1]
    # every trace sees the geometry the sequential loop had after its expand
    ep_idx = np.searchsorted(epochs["first_trace"], np.arange(n), side="right") - 1
    for k in range(n):
        e = epochs[ep_idx[k]]
        xmin, ymin, xmax, ymax, nr, nc = geo[k]
        assert (e["xmin"], e["ymin"], e["nrows"], e["ncols"]) == (xmin, ymin, nr, nc)
        # trace-time node (0, 0) lands on the final node at the same coordinates
        assert abs(final.xmin + e["col_off"] * dx - xmin) < 1e-6 * max(1.0, abs(xmin))
        assert abs(final.ymin + e["row_off"] * dx - ymin) < 1e-6 * max(1.0, abs(ymin))


def test_probabilityfield_api_matches_reference_tests():
    e = golden("reference_kats.json")["expand"]
    with pytest.raises(RangeError):
        ProbabilityField(0, 0)
    pf = ProbabilityField(1, 1)
    assert pf.deltax == 1 and pf.deltay == 1
    with pytest.raises(RangeError):
        pf.expand(1, 0, 1, 0)
    pf.expand(*e["bbox"])
    assert (pf.nrows, pf.ncols) == (e["nrows"], e["ncols"])
    assert (pf.xmin, pf.xmax, pf.ymin, pf.ymax) == (e["xmin"], e["xmax"], e["ymin"], e["ymax"])
    pf.expand(*e["second_bbox"])
    assert (pf.nrows, pf.ncols) == (e["nrows"], e["ncols"])
    assert ProbabilityField.distancesquared(0, 1, 1, 0, 0, 0) == 0.5


def test_probabilityfield_expand_register_match_oracle():
    rng = np.random.default_rng(5)
    pf = ProbabilityField(2.5, 2.5, 100.0, -40.0)
    of = O.OracleField(2.5, 2.5, 100.0, -40.0)
    for _ in range(40):
        c = rng.normal([100.0, -40.0], 60.0)
        h = np.abs(rng.normal(0, 20.0, 2))
        b = (c[0] - h[0], c[0] + h[0], c[1] - h[1], c[1] + h[1])
        pf.expand(*b)
        of.expand(*b)
        assert (pf.nrows, pf.ncols) == of.shape
        assert (pf.xmin, pf.xmax, pf.ymin, pf.ymax) == (of.s.xmin, of.s.xmax, of.s.ymin, of.s.ymax)
    pf.rgrid[3, 4] = True
    pf.register(0.5)
    assert pf.total_weight == 0.5 and pf.pgrid[3, 4] == 0.5 and not pf.rgrid.any()
    pf.rgrid[1, 1] = True
    pf.reset()
    assert not pf.rgrid.any()


def test_start_points_match_reference_loop():
    for P in (1, 7, 100, 1000):
        sp = start_points(2000.0, 2000.0, 0.25, P)
        ref = pyref.start_points(2000.0, 2000.0, 0.25, P)
        assert np.array_equal(sp, np.array(ref, dtype=float))


def test_distribution_vocabulary():
    rng = np.random.default_rng(1)
    assert generate_random_variate(3.5, rng) == 3.5
    assert 2.0 <= generate_random_variate((2.0, 3.0), rng) <= 3.0
    assert 1.0 <= generate_random_variate((1.0, 2.0, 4.0), rng) <= 4.0
    with pytest.raises(DistributionError):
        generate_random_variate((1.0, 2.0, 3.0, 4.0), rng)
    assert compute_variate_mean(7.0) == 7.0
    assert compute_variate_mean((2.0, 4.0)) == 3.0
    assert compute_variate_mean((1.0, 2.0, 6.0)) == 3.0
    assert isdistribution(0.3, 0, 1) and isdistribution((0.2, 0.3), 0, 1)
    assert not isdistribution((0.3, 0.2), 0, 1) and not isdistribution("x", 0, 1)
    ln = Lognormal(math.log(10.0), 0.5)
    x = ln.sample(rng, 200000)
    assert abs(np.log(x).mean() - math.log(10.0)) < 0.01
    assert abs(np.log(x).std() - 0.5) < 0.01
    assert abs(x.mean() / ln.mean() - 1.0) < 0.01
    sg = Signed(Lognormal(math.log(800.0), 0.6), 0.1)
    v = sg.sample(rng, 200000)
    assert abs((v < 0).mean() - 0.1) < 0.005
    tri = rng.triangular(300.0, 500.0, 700.0, 200000)
    assert abs(tri.mean() - compute_variate_mean((300.0, 500.0, 700.0))) < 1.0


def test_sites_are_well_posed():
    for s in (sites.config0(), sites.config1(10)):
        wxy = s.wells_xy()
        obs = s.obs_array()
        d = np.hypot(obs[:, None, 0] - wxy[None, :, 0], obs[:, None, 1] - wxy[None, :, 1])
        assert d.min() >= 100.0
        assert (obs[:, 2] > s.base).all() and (obs[:, 3] == 0.1).all()
        pf = s.fixed_field()
        assert (pf.nrows, pf.ncols) == (s.nrows, s.ncols)
        assert pf.xmin < s.xt < pf.xmax and pf.ymin < s.yt < pf.ymax
    a = sites.draw(sites.config0(), seed=3)
    b = sites.draw(sites.config0(), seed=3)
    for k in a:
        assert np.array_equal(a[k], b[k])



def test_qmc_split_keeps_within_the_budget():
    from onekapy_amd.surrogate import qmc_split
    assert qmc_split(500000, 32, pow2=False) == (15625, 32)
    assert qmc_split(500000, 64) == (4096, 122)
    assert qmc_split(500000, 16) == (16384, 30)
    for m in (1, 7, 4096, 10 ** 6 + 3):
        for r in (1, 3, 64):
            if m < r:
                with pytest.raises(ValueError):
                    qmc_split(m, r)
                continue
            npts, nrep = qmc_split(m, r)
            assert npts & (npts - 1) == 0 and nrep >= r and npts * nrep <= m
            assert npts * (nrep + 1) > m          # no whole replicate left out
