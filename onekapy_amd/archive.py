"""Archive an Oneka run as a bzip2 pickle, in the reference's format
(oneka/archive.py:46-89): a dict of the run's inputs plus the
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


class _FieldRecord:
    """Attribute state of a ProbabilityField.  Archives name it by the
    reference's class path, oneka.probabilityfield.ProbabilityField."""


class _ArchivePickler(pickle._Pickler):
    """A pickler that writes the class _FieldRecord as the global
    oneka.probabilityfield.ProbabilityField without looking that module up, so
    writing an archive never imports (or registers) a module named oneka.  The
    class is intercepted in save() itself, ahead of the type dispatch table,
    which other libraries in the process may have extended."""

    def save(self, obj, save_persistent_id=True):
        if obj is not _FieldRecord or id(obj) in self.memo:
            return super().save(obj, save_persistent_id)
        self.framer.commit_frame()
        if self.proto >= 4:
            super().save(REF_MODULE)
            super().save(REF_CLASS)
            self.write(pickle.STACK_GLOBAL)
        else:
            self.write(pickle.GLOBAL + (REF_MODULE + "\n" + REF_CLASS + "\n").encode("ascii"))
        self.memoize(obj)


def _as_reference_field(pf):
    out = _FieldRecord()
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
    buf = io.BytesIO()
    _ArchivePickler(buf, protocol=pickle.DEFAULT_PROTOCOL).dump(oneka_dict)
    with bz2.BZ2File(filename, "w") as fp:
        fp.write(buf.getvalue())
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
