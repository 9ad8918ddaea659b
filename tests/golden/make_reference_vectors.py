"""Write golden vectors by running the OnekaPy reference (development container
only; the reference never travels with this repository).

    python tests/golden/make_reference_vectors.py [REFERENCE_ROOT]

Writes tests/golden/reference_vectors.npz (inputs and outputs only).  These pin
what the other checks leave open: the exact accepted-vertex sequence of
compute_backtrace (step-control constants 0.1, 0.9, 10, 1/5 and the clamp),
rasterization with the expansion history, and a whole deterministic capture
zone.  numpy 2.x removed np.float, which the reference's ProbabilityField uses
(probabilityfield.py:148-149, 215-216, 249-250); it is aliased to float here.
"""
import ast
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")
sys.path.insert(0, REF)
if not hasattr(np, "float"):
    np.float = float

from oneka.capturezone import compute_backtrace, compute_capturezone  # noqa: E402
from oneka.deterministic import create_deterministic_capturezone  # noqa: E402
from oneka.model import Model  # noqa: E402
from oneka.probabilityfield import ProbabilityField  # noqa: E402


def data_module(name):
    """UPPERCASE constants of a reference data file, parsed without importing it."""
    with open(os.path.join(REF, "data", name + ".py")) as f:
        tree = ast.parse(f.read())
    out = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and isinstance(node.targets[0], ast.Name):
            try:
                out[node.targets[0].id] = ast.literal_eval(node.value)
            except ValueError:
                out[node.targets[0].id] = eval(compile(ast.Expression(node.value), name, "eval"),
                                               {"__builtins__": {}})
    return out


def pack_traces(traces):
    lens = np.array([len(t) for t in traces], dtype=np.int64)
    xy = np.array([[float(v[0]), float(v[1])] for t in traces for v in t], dtype=np.float64)
    return lens, xy


def main():
    out = {}
    # 1. backtraces through a 3-well field (one injection well), both velocities
    wells = [[2250.0, 2250.0, 0.25, 750.0], [1750.0, 2750.0, 0.25, 700.0],
             [2600.0, 1900.0, 0.25, -300.0]]
    coef = np.array([-1e-5, -2e-5, 1e-6, 3e-3, -2e-3, 1000.0])
    mo = Model(0.0, 20.0, 0.2, 22.0, wells, 2250.0, 2250.0, coef)

    def fe_conf(xy):
        vx, vy = mo.compute_velocity_confined(xy[0], xy[1])
        return np.array([-vx, -vy])

    def fe_unconf(xy):
        vx, vy = mo.compute_velocity(xy[0], xy[1])
        return np.array([-vx, -vy])

    th = np.linspace(0, 2 * np.pi, 13)[:-1]
    starts = np.array([(1.25 * np.cos(t) + 2250.0, 1.25 * np.sin(t) + 2250.0) for t in th])
    conf = [compute_backtrace(x, y, 3652.5, 1, 10, fe_conf) for x, y in starts]
    unconf = [compute_backtrace(x, y, 1000.0, 1, 20, fe_unconf) for x, y in starts[:4]]
    out["bt_wells"] = np.array(wells)
    out["bt_model"] = np.array([0.0, 20.0, 0.2, 22.0, 2250.0, 2250.0])
    out["bt_coef"] = coef
    out["bt_starts"] = starts
    out["bt_conf_len"], out["bt_conf_xy"] = pack_traces(conf)
    out["bt_unconf_len"], out["bt_unconf_xy"] = pack_traces(unconf)

    # 2. two compute_capturezone calls into one field from its 3x3 start
    pf = ProbabilityField(10, 10, 2250.0, 2250.0)
    coef2 = np.array([2e-5, -1e-5, 0.0, -4e-3, 1e-3, 800.0])
    compute_capturezone(2250.0, 2250.0, 0.25, 24, 3652.5, pf, 20, 1.0, 1, 10, fe_conf)
    mo.coef = coef2
    compute_capturezone(2250.0, 2250.0, 0.25, 16, 3652.5, pf, 20, 0.5, 1, 10, fe_conf)
    out["cz_coef2"] = coef2
    out["cz_geom"] = np.array([pf.xmin, pf.xmax, pf.ymin, pf.ymax, pf.nrows, pf.ncols,
                               pf.total_weight])
    out["cz_pgrid"] = pf.pgrid

    # 3. the deterministic capture zone of data/basic_deterministic.py, coarsened
    d = data_module("basic_deterministic")
    spacing, umbra, npaths = 10, 20, 24
    pf = create_deterministic_capturezone(
        d["TARGET"], npaths, d["DURATION"], d["BASE"], d["C_DIST"], d["P_DIST"], d["T_DIST"],
        d["WELLS"], d["OBSERVATIONS"], spacing, umbra, d["CONFINED"], d["TOL"], d["MAXSTEP"])
    # the fitted coefficients it used (deterministic.py:180-199)
    mean = lambda a: a if type(a) is not tuple else (sum(a) / len(a))
    wl = [[w[0], w[1], w[2], mean(w[3])] for w in d["WELLS"]]
    m2 = Model(d["BASE"], mean(d["C_DIST"]), mean(d["P_DIST"]), mean(d["T_DIST"]), wl)
    ev, cov = m2.fit_regional_flow(d["OBSERVATIONS"], wl[0][0], wl[0][1])
    out["det_params"] = np.array([npaths, d["DURATION"], d["BASE"], mean(d["C_DIST"]),
                                  mean(d["P_DIST"]), mean(d["T_DIST"]), spacing, umbra, d["TOL"],
                                  d["MAXSTEP"]])
    out["det_wells"] = np.array(wl)
    out["det_obs"] = np.array(d["OBSERVATIONS"], dtype=np.float64)
    out["det_ev"] = ev.ravel()
    out["det_cov"] = cov
    out["det_geom"] = np.array([pf.xmin, pf.xmax, pf.ymin, pf.ymax, pf.nrows, pf.ncols,
                                pf.total_weight])
    out["det_pgrid"] = pf.pgrid
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
