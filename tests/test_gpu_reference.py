"""GPU path against vectors written by the reference itself
(tests/golden/make_reference_vectors.py, make_reference_sites.py)."""
import hashlib
import os

import numpy as np
import pytest

from onekapy_amd import (Model, ProbabilityField, Realizations, _lib, chronicle_into_field,
                         compute_backtrace, compute_capturezone, create_deterministic_capturezone,
                         start_points)
from onekapy_amd.capturezone import trace_vertices
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


# Vectors the reference wrote on the benchmark sites
# (tests/golden/make_reference_sites.py).

@pytest.fixture(scope="module")
def sites_ref():
    return np.load(os.path.join(GOLDEN, "reference_sites.npz"))


def site_params(ref, cfg):
    base, xt, yt, rt, npaths, duration, spacing, umbra, tol, maxstep = ref[cfg + "_params"]
    return dict(base=base, xt=xt, yt=yt, rt=rt, npaths=int(npaths), duration=duration,
                spacing=spacing, umbra=umbra, tol=tol, maxstep=maxstep)


def vertex_hash(v):
    return hashlib.sha256(np.ascontiguousarray(v, dtype=np.float64).tobytes()).hexdigest()


def assert_site_field(pf, ref, cfg):
    g = ref[cfg + "_geom"]
    assert (pf.xmin, pf.xmax, pf.ymin, pf.ymax) == tuple(g[:4])
    assert (pf.nrows, pf.ncols) == (int(g[4]), int(g[5])) and pf.total_weight == g[6]
    pg = pf.pgrid.ravel()
    nz = np.flatnonzero(pg)
    assert np.array_equal(nz, ref[cfg + "_nodes"])
    assert np.array_equal(pg[nz], ref[cfg + "_values"])


def expected_status(ref, cfg):
    """0 for each trace, 1 (truncated) for those the reference ends dry."""
    pick = ref[cfg + "_trace_index"]
    dry = ref[cfg + "_dry"] if cfg + "_dry" in ref.files else []
    return np.isin(pick, dry).astype(np.int32)


def assert_site_traces(wells, aquifer, coef, p, starts, ref, cfg, confined=True):
    real = Realizations(aquifer[0], aquifer[1], aquifer[2], wells[:, 3], coef)
    verts, nverts, status = trace_vertices(wells[:, :2], p["xt"], p["yt"], real, starts,
                                           p["duration"], p["tol"], p["maxstep"], confined)
    assert np.array_equal(status, expected_status(ref, cfg))
    assert np.array_equal(nverts, ref[cfg + "_trace_len"])
    for k in range(len(starts)):
        v = verts[k, :nverts[k]]
        assert np.array_equal(v[-1], ref[cfg + "_trace_last"][k])
        assert vertex_hash(v) == ref[cfg + "_trace_sha256"][k]


def test_config1_two_realizations_equal_reference(gpu_ctx, sites_ref):
    """2,000 traces of the 25-well site through the reference's own API
    (compute_capturezone twice into one field, weights 1 and 0.5), and the
    vertex hashes of every 25th trace of the first realization."""
    ref, p = sites_ref, site_params(sites_ref, "cfg1")
    pf = ProbabilityField(p["spacing"], p["spacing"], p["xt"], p["yt"])
    for r, weight in enumerate(ref["cfg1_weights"]):
        K, n, b = ref["cfg1_aquifer"][r]
        m = Model(p["base"], K, n, b, [tuple(w) for w in ref["cfg1_wells"][r]], p["xt"], p["yt"],
                  ref["cfg1_coef"][r])
        compute_capturezone(p["xt"], p["yt"], p["rt"], p["npaths"], p["duration"], pf,
                            p["umbra"], weight, p["tol"], p["maxstep"], m.velocity_function(True))
    assert_site_field(pf, ref, "cfg1")
    starts = start_points(p["xt"], p["yt"], p["rt"], p["npaths"])[ref["cfg1_trace_index"]]
    assert_site_traces(ref["cfg1_wells"][0], ref["cfg1_aquifer"][0], ref["cfg1_coef"][0], p,
                       starts, ref, "cfg1")


@pytest.mark.parametrize("schedule", [_lib.SCHED_WORKGROUP, _lib.SCHED_POOL],
                         ids=["workgroup", "pool"])
@pytest.mark.parametrize("cfg", ["cfg4", "cfg4long", "cfg2s2", "cfg2s50", "cfg1u", "cfg0dry"])
def test_site_traces_equal_reference(gpu_ctx, sites_ref, cfg, schedule):
    """32 traces on each of the 200-well site (injection wells included) and
    two sites of the uneven 64-site workload (135 and 18 wells), 32 traces of
    the unconfined velocity on the 25-well site, four long traces (2,701 to
    9,077 vertices) on the 200-well site, and unconfined traces dry at their
    start point beside wet ones: the FP64-exact
    trace-and-rasterize kernel on either schedule registers the reference's
    field, and the vertex sequences hash to the reference's."""
    ref, p = sites_ref, site_params(sites_ref, cfg)
    confined = bool(ref[cfg + "_confined"][0])
    wells = ref[cfg + "_wells"]
    starts = start_points(p["xt"], p["yt"], p["rt"], p["npaths"])[ref[cfg + "_trace_index"]]
    K, n, b = ref[cfg + "_aquifer"]
    real = Realizations([K], [n], [b], wells[None, :, 3], ref[cfg + "_coef"][None])
    pf = ProbabilityField(p["spacing"], p["spacing"], p["xt"], p["yt"])
    res = chronicle_into_field(pf, wells[:, :2], p["xt"], p["yt"], real, starts, p["duration"],
                               p["tol"], p["maxstep"], p["umbra"], confined, _lib.FP64_EXACT,
                               schedule=schedule)
    assert np.array_equal(res.trace_status, expected_status(ref, cfg))
    pf.rgrid |= res.counts > 0
    pf.register(1.0)
    assert_site_field(pf, ref, cfg)
    assert_site_traces(wells, ref[cfg + "_aquifer"], ref[cfg + "_coef"], p, starts, ref, cfg,
                       confined)
