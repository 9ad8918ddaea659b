"""Deterministic capture zone (oneka/deterministic.py:65-235) on the GPU.

The nrealizations == 1 path of the reference: every distribution is replaced by
its mean, the regional flow is fitted once and its expected coefficients used
as they are (no MVN draw), and one capture zone is registered with weight 1.0.
"""
import logging

import numpy as np

from .capturezone import compute_capturezone
from .model import Model
from .probabilityfield import ProbabilityField
from .stochastic import compute_variate_mean

log = logging.getLogger("Oneka")


def create_deterministic_capturezone(target, npaths, duration, base, c_dist, p_dist, t_dist,
                                     stochastic_wells, observations, spacing, umbra, confined,
                                     tol, maxstep, precision="fp64", max_steps=0):
    """Compute the deterministic capture zone for the target well (deterministic.py:65-235)."""
    xtarget, ytarget, rtarget = stochastic_wells[target][0:3]
    pfield = ProbabilityField(spacing, spacing, xtarget, ytarget)
    wells = []
    for w in stochastic_wells:
        xw, yw, rw = w[0:3]
        wells.append([xw, yw, rw, compute_variate_mean(w[3])])
    conductivity = compute_variate_mean(c_dist)
    porosity = compute_variate_mean(p_dist)
    thickness = compute_variate_mean(t_dist)
    mo = Model(base, conductivity, porosity, thickness, wells)
    coef_ev, coef_cov = mo.fit_regional_flow(observations, xtarget, ytarget)
    mo.coef = np.reshape(coef_ev, [6, ])
    if log.isEnabledFor(logging.INFO):
        log.info("\n")
        log.info("Deterministic Capture Zone Parameters: ")
        log.info(f"base         = {base:.2f}")
        log.info(f"conductivity = {conductivity:.2f}")
        log.info(f"porosity     = {porosity:.2f}")
        recharge = -2 * (mo.coef[0] + mo.coef[1])
        log.info(f"recharge     = {recharge:+.4e}")
        log.info("\n")
        log.info("Conic Discharge Potential Coefficients: ")
        for name, value in zip("ABCDEF", mo.coef):
            log.info(f"{name} = {value:+.4e}")
    compute_capturezone(xtarget, ytarget, rtarget, npaths, duration, pfield, umbra, 1.0, tol,
                        maxstep, mo.velocity_function(confined), precision=precision,
                        max_steps=max_steps)
    return pfield
