"""GPU path against vectors written by the reference itself
(tests/golden/make_reference_vectors.py)."""
import os

import numpy as np
import pytest

from onekapy_amd import (Model, ProbabilityField, compute_backtrace, compute_capturezone,
                         create_deterministic_capturezone)
from helpers import GOLDEN

pytestmark = pytest.mark.gpu


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
    return Model(base, K, n, b, [tuple(w) for w in ref["bt_wells"]], xo, yo,
                 ref["bt_coef"] if coef is None else coef)


def test_backtrace_vertices_equal_reference(gpu_ctx, ref):
    m = bt_model(ref)
    conf = m.velocity_function(True)
    for (xs, ys), want in zip(ref["bt_starts"], unpack(ref["bt_conf_len"], ref["bt_conf_xy"])):
        got = np.array(compute_backtrace(xs, ys, 3652.5, 1, 10, conf))
        assert np.array_equal(got, want)
    unconf = m.velocity_function(False)
    for (xs, ys), want in zip(ref["bt_starts"][:4],
                              unpack(ref["bt_unconf_len"], ref["bt_unconf_xy"])):
        got = np.array(compute_backtrace(xs, ys, 1000.0, 1, 20, unconf))
        assert np.array_equal(got, want)


def test_capturezone_sequence_equals_reference(gpu_ctx, ref):
    pf = ProbabilityField(10, 10, 2250.0, 2250.0)
    compute_capturezone(2250.0, 2250.0, 0.25, 24, 3652.5, pf, 20, 1.0, 1, 10,
                        bt_model(ref).velocity_function(True))
    compute_capturezone(2250.0, 2250.0, 0.25, 16, 3652.5, pf, 20, 0.5, 1, 10,
                        bt_model(ref, ref["cz_coef2"]).velocity_function(True))
    g = ref["cz_geom"]
    assert (pf.xmin, pf.xmax, pf.ymin, pf.ymax) == tuple(g[:4])
    assert (pf.nrows, pf.ncols) == (int(g[4]), int(g[5])) and pf.total_weight == g[6]
    assert np.array_equal(pf.pgrid, ref["cz_pgrid"])


def test_deterministic_capturezone_equals_reference(gpu_ctx, ref):
    npaths, duration, base, K, n, b, spacing, umbra, tol, maxstep = ref["det_params"]
    wells = [tuple(w) for w in ref["det_wells"]]
    obs = [tuple(o) for o in ref["det_obs"]]
    pf = create_deterministic_capturezone(0, int(npaths), duration, base, K, n, b, wells, obs,
                                          spacing, umbra, True, tol, maxstep)
    g = ref["det_geom"]
    assert (pf.xmin, pf.xmax, pf.ymin, pf.ymax) == tuple(g[:4])
    assert (pf.nrows, pf.ncols) == (int(g[4]), int(g[5]))
    assert np.array_equal(pf.pgrid, ref["det_pgrid"])
