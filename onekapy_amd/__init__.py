"""onekapy_amd -- OnekaPy's Monte Carlo capture-zone hot path on AMD MI355X.

The public API mirrors the reference's modules:
  Model                          (oneka/model.py)
  compute_capturezone,
  compute_backtrace              (oneka/capturezone.py)
  ProbabilityField               (oneka/probabilityfield.py)
  create_stochastic_capturezone,
  generate_random_variate,
  compute_variate_mean,
  isdistribution                 (oneka/stochastic.py)
  create_deterministic_capturezone (oneka/deterministic.py)
  filter_obs                     (oneka/utilities.py)
  dump_oneka, load_oneka         (oneka/archive.py)
  impact_curve, decile_table     (oneka/visualize.py create_impact_plot, oneka.py deciles)
and adds SiteBatch (many sites in one graph-captured GPU run) with its
autotuner, and the learned control variate (surrogate.py).
All evaluation, fitting, tracing and rasterization run in liboka.so
(hand-written HIP for gfx950, C ABI in include/oka.h) through ctypes.
"""
from .model import Model, AquiferError, RangeError
from .probabilityfield import ProbabilityField
from .capturezone import (ModelVelocity, compute_backtrace, compute_capturezone, start_points,
                          trace_vertices)
from .stochastic import (DistributionError, Lognormal, Signed, compute_variate_mean,
                         create_stochastic_capturezone, draw_realizations, fit_and_draw,
                         generate_random_variate, isdistribution)
from .deterministic import create_deterministic_capturezone
from .engine import Realizations, chronicle_into_field
from .utilities import filter_obs, validate_oneka_inputs
from .archive import dump_oneka, load_oneka
from .impact import decile_table, impact_curve
from .batch import SiteBatch
from . import _lib

__all__ = [
    "Model", "AquiferError", "RangeError", "ProbabilityField", "ModelVelocity",
    "compute_backtrace", "compute_capturezone", "start_points", "trace_vertices",
    "DistributionError", "Lognormal", "Signed", "compute_variate_mean",
    "create_stochastic_capturezone", "create_deterministic_capturezone", "draw_realizations",
    "fit_and_draw",
    "generate_random_variate", "isdistribution", "Realizations", "chronicle_into_field",
    "filter_obs", "validate_oneka_inputs", "dump_oneka", "load_oneka", "impact_curve",
    "decile_table", "SiteBatch",
]
