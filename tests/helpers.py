"""Shared test helpers: run a site through the oracle with the same inputs the
GPU path gets."""
import json
import os

import numpy as np

import oracle as O
from onekapy_amd import sites
from onekapy_amd.capturezone import start_points

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def oracle_fit_draw(site, d, math=O.MATH_CR):
    """Oracle fit + MVN draw for every realization of draws `d`."""
    wxy = site.wells_xy()
    R = len(d["K"])
    ev = np.zeros((R, 6))
    cov = np.zeros((R, 6, 6))
    coef = np.zeros((R, 6))
    for r in range(R):
        m = O.OracleModel(site.base, d["K"][r], d["n"][r], d["b"][r],
                          np.c_[wxy, np.full(len(wxy), 0.25), d["q"][r]])
        ev[r], cov[r] = m.fit(site.obs_array(), site.xt, site.yt, math)
        coef[r] = O.mvn(ev[r], cov[r], d["z"][r])
    return ev, cov, coef


def oracle_grid(pf):
    return O.make_grid(pf.xmin, pf.ymin, pf.xmax, pf.ymax, pf.deltax, pf.deltay, pf.nrows,
                       pf.ncols)


def oracle_chronicle(site, d, coef, pf, npaths=None, math=O.MATH_CR, confined=True,
                     nthreads=0):
    P = site.npaths if npaths is None else npaths
    starts = start_points(site.xt, site.yt, site.rt, P)
    opts = O.trace_opts(site.duration, site.tol, site.maxstep, confined, math)
    return O.chronicle_fixed(site.wells_xy(), site.xt, site.yt, site.base, d["K"], d["n"], d["b"],
                             d["q"], coef, starts, site.umbra, opts, oracle_grid(pf),
                             nthreads=nthreads)


def small_site(nreal=8, npaths=32, seed=0):
    s = sites.config0()
    s.nrealizations = nreal
    s.npaths = npaths
    return s, sites.draw(s, seed=seed)
