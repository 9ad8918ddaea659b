"""Site preprocessing of the reference's driver: observation filtering
(oneka/utilities.py:34-144) and the argument checks of oneka()
(oneka/oneka.py:207-242).  Host-side, once per site."""
import logging

import numpy as np

from .stochastic import isdistribution

log = logging.getLogger("Oneka")

TOO_CLOSE = 1.0   # observations closer than this [m] are one location (utilities.py:98)


def filter_obs(observations, wellfield, buffer):
    """Drop observations within `buffer` [m] of any well; merge each run of
    consecutive observations lying within 1 m of the run's first one into
    their minimum-variance weighted mean, with standard deviation
    sqrt(1 / sum 1/sigma_k^2) (utilities.py:101-131).  Returns the list of
    retained (x, y, z_ev, z_std) tuples, merged ones included."""
    obs = []
    for ob in observations:
        near = next((we for we in wellfield
                     if np.hypot(ob[0] - we[0], ob[1] - we[1]) <= buffer), None)
        if near is None:
            obs.append(ob)
        else:
            log.info('observation removed: {0} is too close to {1}'.format(ob, near))
    kept = []
    i = 0
    while i < len(obs):
        j = i + 1
        while j < len(obs) and np.hypot(obs[i][0] - obs[j][0], obs[i][1] - obs[j][1]) < TOO_CLOSE:
            j += 1
        if j - i > 1:
            num = 0
            den = 0
            for k in range(i, j):
                num += obs[k][2] / obs[k][3] ** 2
                den += 1 / obs[k][3] ** 2
                log.info('    duplicate observation: {0}'.format(obs[k]))
            kept.append((obs[i][0], obs[i][1], num / den, np.sqrt(1 / den)))
            log.info('averaged observation created: {0}'.format(kept[-1]))
        else:
            kept.append(obs[i])
        i = j
    log.info('active observations: {0}'.format(len(kept)))
    return kept


def _number(v):
    return isinstance(v, int) or isinstance(v, float)


def validate_oneka_inputs(projectname, target, npaths, duration, base, c_dist, p_dist, t_dist,
                          stochastic_wells, observations, buffer, spacing, umbra, smooth,
                          confined, tol, maxstep, runtime=""):
    """The checks oneka() makes before a run (oneka.py:207-242); raises
    AssertionError on the first that fails, as the reference does."""
    wells = stochastic_wells if isinstance(stochastic_wells, list) else []
    checks = [
        (isinstance(projectname, str), "projectname is a str"),
        (isinstance(runtime, str), "runtime is a str"),
        (isinstance(target, int) and 0 <= target < len(wells), "target indexes stochastic_wells"),
        (isinstance(npaths, int) and 0 < npaths, "npaths is a positive int"),
        (_number(duration) and 0 != duration, "duration is a nonzero number"),
        (_number(base), "base is a number"),
        (isdistribution(c_dist, 0, np.inf), "c_dist is a distribution on [0, inf]"),
        (isdistribution(p_dist, 0, 1), "p_dist is a distribution on [0, 1]"),
        (isdistribution(t_dist, 0, np.inf), "t_dist is a distribution on [0, inf]"),
        (isinstance(stochastic_wells, list) and len(stochastic_wells) >= 1,
         "stochastic_wells is a non-empty list"),
    ]
    for ok, what in checks:
        if not ok:
            raise AssertionError(what)
    for we in stochastic_wells:
        if not (len(we) == 4 and _number(we[0]) and _number(we[1]) and _number(we[2])
                and 0 < we[2] and isdistribution(we[3], -np.inf, np.inf)):
            raise AssertionError(f"bad well {we!r}")
    if not (isinstance(observations, list) and len(observations) > 6):
        raise AssertionError("observations is a list of more than 6")
    for ob in observations:
        if not (len(ob) == 4 and all(_number(v) for v in ob) and 0 <= ob[3]):
            raise AssertionError(f"bad observation {ob!r}")
    for name, v in (("buffer", buffer), ("spacing", spacing), ("umbra", umbra),
                    ("smooth", smooth)):
        if not (_number(v) and 0 < v):
            raise AssertionError(f"{name} is a positive number")
    if not isinstance(confined, bool):
        raise AssertionError("confined is a bool")
    for name, v in (("tol", tol), ("maxstep", maxstep)):
        if not (_number(v) and 0 < v):
            raise AssertionError(f"{name} is a positive number")
