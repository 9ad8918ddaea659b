"""Pins for the CPU oracle (no GPU): the reference's own known answers, library
routines, closed forms, brute force and exact arithmetic.  Each check is chosen
so that a plausible transcription mistake (dropped term, wrong sign or index,
transposed operand, wrong constant) fails it."""
import math
from decimal import Decimal, getcontext
from fractions import Fraction as F

import numpy as np
import pytest

import oracle as O
from oracle import pyref
from helpers import golden

MATHS = (O.MATH_GLIBC, O.MATH_CR)


def kat_model():
    k = golden("reference_kats.json")["model"]
    return O.OracleModel(k["base"], k["conductivity"], k["porosity"], k["thickness"], k["wells"],
                         k["xo"], k["yo"], k["coef"])


@pytest.mark.parametrize("math_mode", MATHS)
def test_model_kats(oracle_lib, math_mode):
    g = golden("reference_kats.json")
    m = kat_model()
    assert np.isclose(m.potential(*g["potential"]["xy"], math=math_mode), g["potential"]["value"],
                      rtol=1e-6)
    assert np.isclose(m.head(*g["head"]["xy"], math=math_mode), g["head"]["value"], rtol=1e-6)
    assert np.allclose(m.discharge(*g["discharge"]["xy"]), g["discharge"]["value"], rtol=1e-6)
    for case in g["velocity"]["cases"]:
        assert np.allclose(m.velocity(*case["xy"], math=math_mode), case["value"], rtol=1e-6)
        # head > thickness at both points: the confined kludge gives the same V
        assert np.allclose(m.velocity_confined(*case["xy"]), case["value"], rtol=1e-6)


def test_discharge_is_minus_gradient_of_potential(oracle_lib):
    """Q = -grad(Phi): a wrong sign or a dropped regional term fails this."""
    m = kat_model()
    m.set_coef([2e-4, -3e-4, 1.5e-4, 0.3, -0.2, 500.0])
    h = 1e-3
    for x, y in [(37.0, 91.0), (250.0, -40.0), (-120.0, 310.0)]:
        gx = (m.potential(x + h, y) - m.potential(x - h, y)) / (2 * h)
        gy = (m.potential(x, y + h) - m.potential(x, y - h)) / (2 * h)
        assert np.allclose(m.discharge(x, y), [-gx, -gy], rtol=1e-6, atol=1e-6)


def test_single_well_radial_velocity_closed_form(oracle_lib):
    q, b, n = 600.0, 20.0, 0.25
    m = O.OracleModel(0.0, 10.0, n, b, [(0.0, 0.0, 0.25, q)])
    for x, y in [(3.0, 4.0), (-10.0, 2.0), (100.0, -250.0)]:
        r = math.hypot(x, y)
        v = q / (2 * math.pi * r * b * n)           # towards the well (pumping)
        assert np.allclose(m.velocity_confined(x, y), [-v * x / r, -v * y / r], rtol=1e-13)


@pytest.mark.parametrize("math_mode", MATHS)
def test_fit_kat_and_library(oracle_lib, math_mode):
    g = golden("reference_kats.json")["fit"]
    m = kat_model()
    ev, cov = m.fit(g["obs"], *g["origin"], math=math_mode)
    assert np.allclose(ev, g["ev"], rtol=g["rtol"])
    assert np.allclose(cov, g["cov"], rtol=g["rtol"])
    # numpy's lstsq (SVD) and inv on the same weighted rows (model.py:598-604)
    k = golden("reference_kats.json")["model"]
    pm = pyref.PyModel(k["base"], k["conductivity"], k["porosity"], k["thickness"], k["wells"])
    ev_np, cov_np = pm.fit([tuple(o) for o in g["obs"]], *g["origin"])
    assert np.allclose(ev, ev_np.ravel(), rtol=1e-10, atol=0)
    assert np.allclose(cov, cov_np, rtol=1e-9, atol=0)


def test_fit_recovers_noise_free_coefficients(oracle_lib):
    """Heads generated from known A..F (confined regime) are fitted back."""
    coef = np.array([-2e-5, -1e-5, 5e-6, 1e-2, -2e-2, 3000.0])
    K, b = 10.0, 20.0
    wells = [(400.0, 300.0, 0.25, 800.0), (-300.0, 500.0, 0.25, 300.0)]
    truth = O.OracleModel(0.0, K, 0.25, b, wells, 0.0, 0.0, coef)
    rng = np.random.default_rng(4)
    obs = []
    for x, y in rng.uniform(-1500, 1500, (30, 2)):
        phi = truth.potential(x, y)
        h = (phi + 0.5 * K * b * b) / (K * b)      # head >= b branch
        assert h >= b
        obs.append((x, y, h, 0.1))
    fit = O.OracleModel(0.0, K, 0.25, b, wells)
    ev, _ = fit.fit(obs, 0.0, 0.0)
    assert np.allclose(ev, coef, rtol=1e-7)


def test_fit_range_error(oracle_lib):
    m = O.OracleModel(100.0, 1.0, 0.25, 10.0, [(0.0, 0.0, 1.0, 10.0)])
    obs = [(float(i), float(i * i % 7), 150.0, 0.1) for i in range(1, 9)] + [(50.0, 3.0, 90.0, 0.1)]
    with pytest.raises(ArithmeticError):
        m.fit(obs, 0.0, 0.0)


def test_mvn_factor_and_statistics(oracle_lib):
    g = golden("reference_kats.json")["fit"]
    m = kat_model()
    ev, cov = m.fit(g["obs"], *g["origin"])
    rng = np.random.default_rng(0)
    z = rng.standard_normal((20000, 6))
    draws = np.array([O.mvn(ev, cov, zz) for zz in z])
    # coef = ev + L z is linear in z: the empirical moments match (ev, cov)
    L = np.linalg.cholesky(cov)
    assert np.allclose(draws, ev + z @ L.T, rtol=1e-12, atol=0)
    assert np.allclose(draws.mean(0), ev, atol=4 * np.sqrt(np.diag(cov) / len(z)))
    # whitened sample covariance is the identity within sampling error
    Li = np.linalg.inv(L)
    W = Li @ np.cov(draws.T) @ Li.T
    assert np.abs(W - np.eye(6)).max() < 5.0 / np.sqrt(len(z)) * 2


def test_dormand_prince_tableau_is_textbook():
    """DP5(4) of Dormand & Prince (1980): rows sum to the nodes c, the error
    weights are b - b* of the textbook 5th/4th-order weights, the 7th stage
    (FSAL) carries e7 = -1/40 in the slot the reference reuses for k2."""
    c = [F(1, 5), F(3, 10), F(4, 5), F(8, 9), F(1), F(1)]
    rows = pyref._DP_ROWS + (pyref._DP_B[:6],)
    for row, ci in zip(rows, c):
        assert math.isclose(sum(row), float(ci), rel_tol=1e-15)
    b = [F(35, 384), F(0), F(500, 1113), F(125, 192), F(-2187, 6784), F(11, 84), F(0)]
    bs = [F(5179, 57600), F(0), F(7571, 16695), F(393, 640), F(-92097, 339200), F(187, 2100),
          F(1, 40)]
    e = [bi - bsi for bi, bsi in zip(b, bs)]
    ref_e = [e[0], e[6], e[2], e[3], e[4], e[5]]   # k2 holds the 7th stage
    assert list(pyref._DP_E) == [float(v) for v in ref_e]
    assert list(pyref._DP_B) == [float(v) for v in b[:6]]


def test_backtrace_uniform_flow_is_linear(oracle_lib):
    D, E, b, n = 0.3, -0.2, 10.0, 0.25
    m = O.OracleModel(0.0, 1.0, n, b, np.zeros((0, 4)), 0.0, 0.0, [0, 0, 0, D, E, 0])
    T = 1000.0
    v, res = m.backtrace(5.0, 7.0, T, 1.0, 10.0)
    assert res.status == O.TRACE_OK and res.t_end == T
    assert np.allclose(v[-1], [5.0 + D * T / (b * n), 7.0 + E * T / (b * n)], rtol=1e-12)
    steps = np.hypot(*np.diff(v, axis=0).T)
    assert (steps < 10.0).all()


def test_backtrace_single_well_radial_closed_form(oracle_lib):
    q, b, n, T = 500.0, 25.0, 0.25, 3652.5
    m = O.OracleModel(0.0, 10.0, n, b, [(0.0, 0.0, 0.25, q)])
    r0 = 1.25
    for th in (0.3, 2.0, 4.4):
        v, res = m.backtrace(r0 * math.cos(th), r0 * math.sin(th), T, 1e-7, 10.0)
        r_exact = math.sqrt(r0 * r0 + q * T / (math.pi * b * n))
        r = math.hypot(*v[-1])
        assert res.status == O.TRACE_OK
        assert abs(r - r_exact) < 1e-5 * r_exact
        assert abs(math.atan2(v[-1][1], v[-1][0]) - math.atan2(math.sin(th), math.cos(th))) < 1e-9
    # at the reference tolerances (tol 1 m, maxstep 10 m) every step stays < maxstep
    v, res = m.backtrace(r0, 0.0, T, 1.0, 10.0)
    assert abs(math.hypot(*v[-1]) - r_exact) < 0.05 * r_exact
    assert (np.hypot(*np.diff(v, axis=0).T) < 10.0).all()


def test_backtrace_exponential_regional_flow(oracle_lib):
    A, b, n, T = 2e-4, 10.0, 0.2, 500.0
    m = O.OracleModel(0.0, 1.0, n, b, np.zeros((0, 4)), 0.0, 0.0, [A, A, 0, 0, 0, 0])
    v, res = m.backtrace(3.0, -2.0, T, 1e-8, 10.0)
    grow = math.exp(2 * A * T / (b * n))
    assert np.allclose(v[-1], [3.0 * grow, -2.0 * grow], rtol=1e-7)


def test_backtrace_deviations_are_reported(oracle_lib):
    m = O.OracleModel(0.0, 10.0, 0.25, 20.0, [(0.0, 0.0, 0.25, 500.0), (30.0, 0.0, 0.25, 200.0)])
    v, res = m.backtrace(30.0, 0.0, 1000.0, 1.0, 10.0)     # start on a well: 0/0
    assert res.status == O.TRACE_STALLED and res.nverts == 1
    v, res = m.backtrace(1.25, 0.0, 1000.0, 1.0, 10.0, max_steps=5)
    assert res.status == O.TRACE_STEP_CAP and res.attempted == 5
    # unconfined with a dry aquifer: AquiferError truncates the trace
    dry = O.OracleModel(0.0, 10.0, 0.25, 20.0, [(0.0, 0.0, 0.25, 500.0)], 0.0, 0.0,
                        [0, 0, 0, 0, 0, -1e9])
    v, res = dry.backtrace(1.25, 0.0, 1000.0, 1.0, 10.0, confined=False)
    assert res.status == O.TRACE_TRUNCATED and res.nverts == 1


def test_python_and_c_restatements_agree_bit_for_bit(oracle_lib):
    wells = [(2250.0, 2250.0, 0.25, 750.0), (1750.0, 2750.0, 0.25, 700.0),
             (2600.0, 1900.0, 0.25, -300.0)]
    coef = [-1e-5, -2e-5, 1e-6, 3e-3, -2e-3, 1000.0]
    pm = pyref.PyModel(0.0, 20.0, 0.2, 22.0, wells, 2250.0, 2250.0, coef)
    om = O.OracleModel(0.0, 20.0, 0.2, 22.0, wells, 2250.0, 2250.0, coef)
    starts = pyref.start_points(2250.0, 2250.0, 0.25, 12)

    def fe(xy):
        vx, vy = pm.velocity_confined(xy[0], xy[1])
        return np.array([-vx, -vy])

    for xs, ys in starts:
        v, att, acc, st = pyref.backtrace(xs, ys, 3652.5, 1, 10, fe)
        ov, res = om.backtrace(float(xs), float(ys), 3652.5, 1, 10, math=O.MATH_GLIBC)
        assert np.array_equal(np.array([(float(a), float(c)) for a, c in v]), ov)
        assert (att, acc) == (res.attempted, res.accepted)
    pf = pyref.PyField(10.0, 10.0, 2250.0, 2250.0)
    of = O.OracleField(10.0, 10.0, 2250.0, 2250.0)
    pyref.realization(pm, starts, 3652.5, 1, 10, pf, 20.0)
    of.capturezone(om, np.array(starts, dtype=float), 20.0, 1.0,
                   O.trace_opts(3652.5, 1, 10, True, O.MATH_GLIBC))
    assert (pf.nrows, pf.ncols) == of.shape
    assert np.array_equal(pf.pgrid, of.pgrid)
    assert (pf.xmin, pf.ymin) == (of.s.xmin, of.s.ymin)


def test_expand_kat_and_distance_kat(oracle_lib):
    g = golden("reference_kats.json")
    e = g["expand"]
    f = O.OracleField(e["spacing"], e["spacing"])
    f.expand(*e["bbox"])
    assert f.shape == (e["nrows"], e["ncols"])
    assert (f.s.xmin, f.s.xmax, f.s.ymin, f.s.ymax) == (e["xmin"], e["xmax"], e["ymin"], e["ymax"])
    f.expand(*e["second_bbox"])
    assert f.shape == (e["nrows"], e["ncols"])
    with pytest.raises(ValueError):
        f.expand(1, 0, 1, 0)
    with pytest.raises(ValueError):
        O.OracleField(0, 0)
    assert O.distancesquared(*g["distancesquared"]["args"]) == g["distancesquared"]["value"]


def _exact_d2(ax, ay, bx, by, cx, cy):
    ax, ay, bx, by, cx, cy = map(F, (ax, ay, bx, by, cx, cy))
    vx, vy, wx, wy = bx - ax, by - ay, cx - ax, cy - ay
    l2 = vx * vx + vy * vy
    t = (vx * wx + vy * wy) / l2
    t = min(max(t, F(0)), F(1))
    px, py = ax + t * vx - cx, ay + t * vy - cy
    return px * px + py * py


def test_insert_matches_brute_force(oracle_lib):
    """Every node of a small grid against the exact segment distance: the
    candidate window and the strict umbra test (probabilityfield.py:296-310)."""
    rng = np.random.default_rng(9)
    for trial in range(60):
        dx = float(rng.choice([1.0, 2.5, 10.0]))
        f = O.OracleField(dx, dx, 0.0, 0.0)
        f.expand(-20 * dx, 20 * dx, -15 * dx, 15 * dx)
        a = rng.uniform(-18 * dx, 18 * dx, 2)
        bvec = a + rng.normal(0, 4 * dx, 2)
        umbra = float(rng.uniform(0.3, 3.0) * dx)
        f.insert(a[0], a[1], bvec[0], bvec[1], umbra)
        r = f.rgrid
        for i in range(f.shape[0]):
            for j in range(f.shape[1]):
                cx, cy = f.s.xmin + j * dx, f.s.ymin + i * dx
                d2 = _exact_d2(a[0], a[1], bvec[0], bvec[1], cx, cy)
                inside = d2 < F(umbra) * F(umbra)
                if abs(float(d2) - umbra * umbra) > 1e-9 * umbra * umbra:
                    assert r[i, j] == inside, (trial, i, j)


def test_zero_length_segment_marks_nothing(oracle_lib):
    f = O.OracleField(1.0, 1.0, 0.0, 0.0)
    f.expand(-5, 5, -5, 5)
    f.insert(0.2, 0.3, 0.2, 0.3, 2.0)
    assert not f.rgrid.any()


def test_pow_and_log_modes(oracle_lib):
    getcontext().prec = 60
    rng = np.random.default_rng(2)
    c = Decimal(0.2)
    for x in 10 ** rng.uniform(-4, 16, 400):
        x = float(x)
        assert O.pow02(x, O.MATH_GLIBC) == math.pow(x, 0.2)
        assert O.pow02(x, O.MATH_CR) == float((Decimal(x).ln() * c).exp())
        assert O.log(x, O.MATH_CR) == float(Decimal(x).ln())
        assert O.log(x, O.MATH_GLIBC) == math.log(x)
