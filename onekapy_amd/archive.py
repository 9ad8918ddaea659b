"""Archive an Oneka run as a bzip2 pickle, in the reference's format
(oneka/archive.py:59-94): a dict of the run's inputs plus the
ProbabilityField, with the field stored under the reference's class path
``oneka.probabilityfield.ProbabilityField`` and the reference's attributes,
so that archives move both ways between the reference and this package.

Loading uses a restricted unpickler: it builds only numpy arrays and scalars,
the ProbabilityField (mapped to this package's class) and this package's
distribution types, and refuses every other global, so an archive cannot run
code when it is loaded.
"""
import bz2
import copyreg
import io
import os
import pickle
import sys
import types
from datetime import datetime

import numpy as np

from .probabilityfield import ProbabilityField
from .stochastic import Lognormal, Signed

REF_MODULE, REF_CLASS = "oneka.probabilityfield", "ProbabilityField"
FIELD_STATE = ("deltax", "deltay", "xmin", "xmax", "ymin", "ymax", "nrows", "ncols", "pgrid",
               "rgrid", "total_weight")
KEYS = ("projectname", "runtime", "target", "npaths", "duration", "nrealizations", "base",
        "c_dist", "p_dist", "t_dist", "stochastic_wells", "observations", "buffer", "spacing",
        "umbra", "smooth", "confined", "tol", "maxstep", "pfield")


def _reference_field_class():
    """A class that pickles as oneka.probabilityfield.ProbabilityField.  When the
    reference package itself is importable its class is used; otherwise a
    stand-in holding only the attribute state is registered under that path."""
    try:
        mod = __import__(REF_MODULE, fromlist=[REF_CLASS])
        return getattr(mod, REF_CLASS)
    except ImportError:
        pass
    mod = sys.modules.get(REF_MODULE)
    if mod is None or not hasattr(mod, REF_CLASS):
        pkg = sys.modules.setdefault("oneka", types.ModuleType("oneka"))
        mod = types.ModuleType(REF_MODULE)
        stand_in = type(REF_CLASS, (), {"__module__": REF_MODULE,
                                        "__doc__": "Attribute state of a ProbabilityField."})
        setattr(mod, REF_CLASS, stand_in)
        sys.modules[REF_MODULE] = mod
        pkg.probabilityfield = mod
    return getattr(mod, REF_CLASS)


def _as_reference_field(pf):
    cls = _reference_field_class()
    out = cls.__new__(cls)
    for k in FIELD_STATE:
        setattr(out, k, getattr(pf, k))
    return out


def dump_oneka(projectname, runtime, target, npaths, duration, nrealizations, base, c_dist,
               p_dist, t_dist, stochastic_wells, observations, buffer, spacing, umbra, smooth,
               confined, tol, maxstep, pfield, filename=None):
    """Writes the run to `filename` (default logs/Oneka<YYYYmmddTHHMMSS>.bz2, the
    reference's name in a portable path) and returns the file name."""
    if filename is None:
        filename = os.path.join("logs", "Oneka" + datetime.now().strftime("%Y%m%dT%H%M%S") + ".bz2")
    values = (projectname, runtime, target, npaths, duration, nrealizations, base, c_dist,
              p_dist, t_dist, stochastic_wells, observations, buffer, spacing, umbra, smooth,
              confined, tol, maxstep, _as_reference_field(pfield))
    oneka_dict = dict(zip(KEYS, values))
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    with bz2.BZ2File(filename, "w") as fp:
        pickle.dump(oneka_dict, fp)
    return filename


_ALLOWED = {
    ("numpy", "ndarray"), ("numpy", "dtype"), ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "_reconstruct"), ("numpy.core.multiarray", "scalar"),
    ("numpy._core.multiarray", "scalar"), ("numpy", "float64"), ("numpy", "int64"),
    ("numpy", "bool_"), ("copyreg", "_reconstructor"), ("builtins", "object"),
    ("builtins", "float"), ("builtins", "int"), ("builtins", "tuple"), ("builtins", "list"),
}


class _SafeUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) == (REF_MODULE, REF_CLASS) or (
                module == ProbabilityField.__module__ and name == "ProbabilityField"):
            return ProbabilityField
        if module == "onekapy_amd.stochastic" and name in ("Lognormal", "Signed"):
            return {"Lognormal": Lognormal, "Signed": Signed}[name]
        if (module, name) in _ALLOWED:
            return {("copyreg", "_reconstructor"): copyreg._reconstructor}.get(
                (module, name)) or getattr(__import__(module, fromlist=[name]), name)
        raise pickle.UnpicklingError(f"refusing to load {module}.{name} from an archive")


def load_oneka(filename):
    """The run dict of an archive written by dump_oneka (here or by the
    reference); its 'pfield' is this package's ProbabilityField."""
    with bz2.BZ2File(filename, "r") as fp:
        data = fp.read()
    return _SafeUnpickler(io.BytesIO(data)).load()
