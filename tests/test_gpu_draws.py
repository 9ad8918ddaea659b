"""Device draws (Philox4x32-10) against the Python checker and the target
distributions."""
import math

import numpy as np
import pytest

from oracle import philox
from onekapy_amd import Lognormal, Signed, _lib
from onekapy_amd.stochastic import draw_realizations_device

pytestmark = pytest.mark.gpu


def test_uniform_draws_are_the_philox_stream(gpu_ctx):
    wells = [(0.0, 0.0, 0.25, (0.0, 1.0)), (10.0, 0.0, 0.25, 3.0)]
    seed = 0x1234_5678_9ABC_DEF1
    d = draw_realizations_device(wells, (2.0, 5.0), 0.25, (20.0, 30.0), 64, seed, first=1000)
    for i in range(64):
        r = 1000 + i
        assert d["q"][i, 0] == philox.uniforms(seed, r, 0)[0]
        assert d["q"][i, 1] == 3.0
        assert d["K"][i] == 2.0 + 3.0 * philox.uniforms(seed, r, 2)[0]
        assert d["n"][i] == 0.25
        assert d["b"][i] == 20.0 + 10.0 * philox.uniforms(seed, r, 4)[0]
    # realizations are reproducible one by one, whatever the batch they are in
    d2 = draw_realizations_device(wells, (2.0, 5.0), 0.25, (20.0, 30.0), 8, seed, first=1010)
    assert np.array_equal(d2["K"], d["K"][10:18]) and np.array_equal(d2["z"], d["z"][10:18])


def test_antithetic_pairs_reflect_the_even_draw(gpu_ctx):
    """Realization 2k + 1 takes 2k's uniforms u as 1 - u and its normals
    negated; realization 2k is the plain draw."""
    mu = math.log(800.0)
    wells = [(0.0, 0.0, 0.25, (300.0, 450.0, 700.0)), (1.0, 0.0, 0.25, Signed(Lognormal(mu, 0.6), 0.3))]
    seed, first, n = 0x0DDB_A11, 2000, 256
    args = (wells, (2.0, 5.0), 0.25, (20.0, 30.0), n, seed)
    a = draw_realizations_device(*args, first=first, antithetic=True)
    p = draw_realizations_device(*args, first=first)
    for k in ("q", "K", "n", "b", "z"):
        assert np.array_equal(a[k][0::2], p[k][0::2]), k
    assert np.array_equal(a["z"][1::2], -a["z"][0::2])
    flips = 0
    for i in range(1, n, 2):
        r = first + i - 1
        # uniform: the reflected uniform, bit for bit
        assert a["K"][i] == 2.0 + 3.0 * (1.0 - philox.uniforms(seed, r, 2)[0])
        assert a["b"][i] == 20.0 + 10.0 * (1.0 - philox.uniforms(seed, r, 4)[0])
        # triangular: CDF(v_odd) = 1 - CDF(v_even)
        for v, w in ((a["q"][i, 0], a["q"][i - 1, 0]),):
            def cdf(x):
                return ((x - 300.0) ** 2 / (400.0 * 150.0) if x < 450.0
                        else 1.0 - (700.0 - x) ** 2 / (400.0 * 250.0))
            assert abs(cdf(v) + cdf(w) - 1.0) < 1e-9
        # lognormal: log v_odd + log v_even = 2 mu; the sign variate reflected
        lo, le = a["q"][i, 1], a["q"][i - 1, 1]
        assert abs(math.log(abs(lo)) + math.log(abs(le)) - 2.0 * mu) < 1e-12
        us = philox.uniforms(seed, r, 1, 1)[0]
        assert (le < 0) == (us < 0.3) and (lo < 0) == (1.0 - us < 0.3)
        flips += (lo < 0) != (le < 0)
    assert flips > 0
    with pytest.raises(_lib.OkaError):
        draw_realizations_device(*args[:4], n - 1, seed, first=first, antithetic=True)
    with pytest.raises(_lib.OkaError):
        draw_realizations_device(*args, first=first + 1, antithetic=True)


def test_qmc_draws_are_the_scrambled_sobol_points(gpu_ctx):
    """OKA_QMC: realization j's uniforms are the scrambled Sobol' point j (the
    checker's), dimension by dimension in the documented order."""
    from oracle import sobol
    mu = math.log(800.0)
    wells = [(0.0, 0.0, 0.25, (300.0, 450.0, 700.0)), (1.0, 0.0, 0.25, Signed(Lognormal(mu, 0.6), 0.3)),
             (2.0, 0.0, 0.25, 700.0)]
    seed, first, n = 0x5EED_0F_50B0, 1000, 512
    d = draw_realizations_device(wells, (2.0, 5.0), 0.25, (20.0, 30.0), n, seed, first, qmc=True)
    # dims: K 0, thickness 1 (porosity is constant), normals 2-7, discharges of
    # wells 0 and 1: 8, 9 (well 2 is constant), then well 1's angle 10, sign 11
    v = sobol.direction_numbers(12)
    for i in range(0, n, 7):
        j = first + i
        u = [sobol.uniform(v[dd], j, seed, dd) for dd in range(12)]
        assert d["K"][i] == 2.0 + 3.0 * u[0]
        assert d["b"][i] == 20.0 + 10.0 * u[1] and d["n"][i] == 0.25
        for p in range(3):
            rad = math.sqrt(-2.0 * math.log(u[2 + 2 * p]))
            ang = 2.0 * math.pi * u[3 + 2 * p]
            assert abs(d["z"][i, 2 * p] - rad * math.cos(ang)) < 1e-12
            assert abs(d["z"][i, 2 * p + 1] - rad * math.sin(ang)) < 1e-12
        fc = 150.0 / 400.0
        tri = (300.0 + math.sqrt(u[8] * 400.0 * 150.0) if u[8] < fc
               else 700.0 - math.sqrt((1.0 - u[8]) * 400.0 * 250.0))
        assert abs(d["q"][i, 0] - tri) < 1e-9
        g = math.sqrt(-2.0 * math.log(u[9])) * math.cos(2.0 * math.pi * u[10])
        ln = math.exp(mu + 0.6 * g) * (-1.0 if u[11] < 0.3 else 1.0)
        assert abs(d["q"][i, 1] - ln) < 1e-9 * abs(ln)
        assert d["q"][i, 2] == 700.0
    # 1024 points from 0: K's uniforms fill each interval of width 1/1024 once
    d = draw_realizations_device(wells, (2.0, 5.0), 0.25, (20.0, 30.0), 1024, seed, 0, qmc=True)
    bins = np.floor((d["K"] - 2.0) / 3.0 * 1024).astype(int)
    assert np.array_equal(np.sort(bins), np.arange(1024))
    with pytest.raises(_lib.OkaError):
        draw_realizations_device(wells, (2.0, 5.0), 0.25, (20.0, 30.0), 4, seed, 0, antithetic=True,
                                 qmc=True)
    with pytest.raises(_lib.OkaError):
        draw_realizations_device(wells, (2.0, 5.0), 0.25, (20.0, 30.0), 4, seed, 2 ** 32 - 2,
                                 qmc=True)


def test_distribution_moments(gpu_ctx):
    n = 400000
    wells = [(0.0, 0.0, 0.25, (300.0, 500.0, 700.0)),
             (1.0, 0.0, 0.25, Signed(Lognormal(math.log(800.0), 0.6), 0.1))]
    d = draw_realizations_device(wells, Lognormal(math.log(10.0), 0.5), (0.2, 0.3), (20.0, 30.0),
                                 n, 7)
    tri = d["q"][:, 0]
    assert tri.min() >= 300.0 and tri.max() <= 700.0
    assert abs(tri.mean() - 500.0) < 0.5 and abs(tri.var() - (4e4 * 3) / 18 * 1.0) < 200.0
    sg = d["q"][:, 1]
    assert abs((sg < 0).mean() - 0.1) < 0.003
    la = np.log(np.abs(sg))
    assert abs(la.mean() - math.log(800.0)) < 0.005 and abs(la.std() - 0.6) < 0.005
    lk = np.log(d["K"])
    assert abs(lk.mean() - math.log(10.0)) < 0.005 and abs(lk.std() - 0.5) < 0.005
    assert abs(d["n"].mean() - 0.25) < 0.001 and abs(d["b"].var() - 100.0 / 12.0) < 0.1
    z = d["z"]
    assert np.abs(z.mean(0)).max() < 0.01 and np.abs(z.std(0) - 1.0).max() < 0.01
    assert np.abs(np.corrcoef(z.T) - np.eye(6)).max() < 0.01


def test_mark_moments_are_exact(gpu_ctx):
    from onekapy_amd.surrogate import plain_estimate
    rng = np.random.default_rng(11)
    # 2 row blocks and a ragged last one; a cell count that is not a multiple of 256
    marks = rng.integers(0, 256, (2 * 1024 + 37, 300)).astype(np.uint8)
    marks[:, 0] = 0        # never visited
    marks[:, 1] = 255      # largest mark: the partial sums' range
    p, v = plain_estimate(marks, 1.0)
    m = marks.astype(np.int64)
    n = marks.shape[0]
    assert np.array_equal(p, m.sum(0).astype(np.float64) / n)
    ref_var = (m * m).sum(0) / n - (m.sum(0) / n) ** 2
    assert np.allclose(v * n, ref_var, rtol=1e-12, atol=1e-9)
    assert p[0] == 0.0 and v[0] == 0.0 and p[1] == 255.0 and v[1] == 0.0
    # the sums accumulate; an empty input leaves them untouched
    s1 = np.full(300, 7, np.uint64)
    s2 = np.full(300, 9, np.uint64)
    empty = np.zeros((0, 300), np.uint8)
    gpu_ctx.check(gpu_ctx.lib.oka_mark_moments(gpu_ctx.h, 0, 300, empty.ctypes.data,
                                               s1.ctypes.data, s2.ctypes.data, 0))
    assert (s1 == 7).all() and (s2 == 9).all()
    gpu_ctx.check(gpu_ctx.lib.oka_mark_moments(gpu_ctx.h, n, 300, marks.ctypes.data,
                                               s1.ctypes.data, None, 0))
    assert np.array_equal(s1, m.sum(0).astype(np.uint64) + 7) and (s2 == 9).all()
