"""The oracle against vectors written by the reference itself
(tests/golden/make_reference_vectors.py): exact vertex sequences, the
expansion history of rasterize, and a whole deterministic capture zone."""
import os

import numpy as np
import pytest

import oracle as O
from onekapy_amd.capturezone import start_points
from helpers import GOLDEN


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "reference_vectors.npz"))


def unpack(lens, xy):
    out, k = [], 0
    for n in lens:
        out.append(xy[k:k + n])
        k += n
    return out


def bt_model(ref, coef=None):
    base, K, n, b, xo, yo = ref["bt_model"]
    return O.OracleModel(base, K, n, b, ref["bt_wells"], xo, yo,
                         ref["bt_coef"] if coef is None else coef)


def test_backtrace_vertices_equal_reference(oracle_lib, ref):
    m = bt_model(ref)
    for (xs, ys), want in zip(ref["bt_starts"], unpack(ref["bt_conf_len"], ref["bt_conf_xy"])):
        got, res = m.backtrace(xs, ys, 3652.5, 1, 10, confined=True, math=O.MATH_GLIBC)
        assert np.array_equal(got, want)
        # the correctly rounded pow differs from glibc's in ~0.1 % of steps; the
        # trajectories stay within 1e-9 relative
        cr, _ = m.backtrace(xs, ys, 3652.5, 1, 10, confined=True, math=O.MATH_CR)
        if cr.shape == want.shape:
            assert np.allclose(cr, want, rtol=1e-9, atol=0)
    for (xs, ys), want in zip(ref["bt_starts"][:4],
                              unpack(ref["bt_unconf_len"], ref["bt_unconf_xy"])):
        got, res = m.backtrace(xs, ys, 1000.0, 1, 20, confined=False, math=O.MATH_GLIBC)
        assert np.array_equal(got, want)


def test_capturezone_expansion_history_equals_reference(oracle_lib, ref):
    f = O.OracleField(10.0, 10.0, 2250.0, 2250.0)
    opts = O.trace_opts(3652.5, 1, 10, True, O.MATH_GLIBC)
    f.capturezone(bt_model(ref), start_points(2250.0, 2250.0, 0.25, 24), 20.0, 1.0, opts)
    f.capturezone(bt_model(ref, ref["cz_coef2"]), start_points(2250.0, 2250.0, 0.25, 16), 20.0, 0.5,
                  opts)
    g = ref["cz_geom"]
    assert (f.s.xmin, f.s.xmax, f.s.ymin, f.s.ymax) == tuple(g[:4])
    assert f.shape == (int(g[4]), int(g[5])) and f.s.total_weight == g[6]
    assert np.array_equal(f.pgrid, ref["cz_pgrid"])


def test_deterministic_capturezone_equals_reference(oracle_lib, ref):
    npaths, duration, base, K, n, b, spacing, umbra, tol, maxstep = ref["det_params"]
    wells = ref["det_wells"]
    m = O.OracleModel(base, K, n, b, wells)
    ev, cov = m.fit(ref["det_obs"], wells[0, 0], wells[0, 1], O.MATH_GLIBC)
    # cond(WA^T WA) is ~3e12 here; the normal-equation solve agrees with
    # LAPACK's SVD solve to 1e-9 of each coefficient's own standard error
    se = np.sqrt(np.diag(ref["det_cov"]))
    assert np.all(np.abs(ev - ref["det_ev"]) <= 1e-9 * se + 1e-12 * np.abs(ref["det_ev"]))
    # entries that vanish in exact arithmetic (symmetric observation grid) hold
    # rounding noise only: compare on the correlation scale
    assert np.all(np.abs(cov - ref["det_cov"]) <= 1e-9 * np.outer(se, se))
    # trace with the reference's own fitted coefficients: the field is identical
    m = O.OracleModel(base, K, n, b, wells, wells[0, 0], wells[0, 1], ref["det_ev"])
    f = O.OracleField(spacing, spacing, wells[0, 0], wells[0, 1])
    f.capturezone(m, start_points(wells[0, 0], wells[0, 1], wells[0, 2], int(npaths)), umbra, 1.0,
                  O.trace_opts(duration, tol, maxstep, True, O.MATH_GLIBC))
    g = ref["det_geom"]
    assert (f.s.xmin, f.s.xmax, f.s.ymin, f.s.ymax) == tuple(g[:4])
    assert f.shape == (int(g[4]), int(g[5]))
    assert np.array_equal(f.pgrid, ref["det_pgrid"])
