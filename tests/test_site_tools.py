"""The "next" rows on the host: filter_obs and the oneka() argument checks
against the reference, and the archive format both ways."""
import bz2
import io
import json
import os
import pickle
import pickletools
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from onekapy_amd import ProbabilityField
from onekapy_amd.archive import dump_oneka, load_oneka
from onekapy_amd.utilities import filter_obs, validate_oneka_inputs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_filter_obs_matches_reference():
    with open(os.path.join(GOLDEN, "reference_aux.json")) as f:
        cases = json.load(f)["filter_obs"]
    for case in cases:
        obs = [tuple(o) for o in case["observations"]]
        wells = [tuple(w[:2]) + tuple(w[2:]) for w in case["wellfield"]]
        got = filter_obs(obs, wells, case["buffer"])
        assert [list(map(float, g)) for g in got] == case["retained"]


def _valid():
    wells = [[0.0, 0.0, 0.25, (300.0, 500.0, 700.0)], [500.0, 0.0, 0.25, 400.0]]
    obs = [[100.0 * i, 50.0, 10.0, 0.1] for i in range(8)]
    return dict(projectname="p", target=0, npaths=10, duration=3652.5, base=0.0,
                c_dist=(5.0, 15.0), p_dist=(0.2, 0.3), t_dist=(20.0, 30.0),
                stochastic_wells=wells, observations=obs, buffer=100.0, spacing=20.0,
                umbra=40.0, smooth=2.0, confined=True, tol=1.0, maxstep=10.0)


@pytest.mark.parametrize("key,bad", [
    ("target", 2), ("npaths", 0), ("duration", 0), ("c_dist", (15.0, 5.0)),
    ("p_dist", (0.2, 1.5)), ("observations", [[0.0, 0.0, 1.0, 0.1]] * 6),
    ("buffer", 0.0), ("smooth", -1.0), ("confined", 1), ("maxstep", 0.0),
    ("stochastic_wells", [[0.0, 0.0, 0.0, 500.0]])])
def test_validation_rejects_what_oneka_rejects(key, bad):
    validate_oneka_inputs(**_valid())
    args = _valid()
    args[key] = bad
    with pytest.raises(AssertionError):
        validate_oneka_inputs(**args)


def test_reference_archive_loads_into_this_package():
    run = load_oneka(os.path.join(GOLDEN, "reference_archive.bz2"))
    pf = run["pfield"]
    assert isinstance(pf, ProbabilityField)
    ref = np.load(os.path.join(GOLDEN, "reference_impact.npz"))
    assert np.array_equal(pf.pgrid, ref["counts"]) and pf.total_weight == 37.0
    assert (pf.nrows, pf.ncols) == ref["counts"].shape and pf.rgrid.dtype == bool
    assert run["npaths"] == 100 and run["t_dist"] == (20.0, 30.0) and run["confined"] is True
    # the field works: register/expand are this package's methods
    pf.expand(pf.xmin - 100.0, pf.xmax, pf.ymin, pf.ymax)
    assert pf.pgrid.sum() == ref["counts"].sum()


def _globals(path):
    with bz2.BZ2File(path) as fp:
        data = fp.read()
    names, strings = set(), []
    for op, arg, _ in pickletools.genops(data):
        if op.name in ("SHORT_BINUNICODE", "BINUNICODE", "UNICODE"):
            strings.append(arg)
        if op.name == "STACK_GLOBAL" and len(strings) >= 2:
            names.add((strings[-2], strings[-1]))
        if op.name == "GLOBAL":
            names.add(tuple(arg.split(" ", 1)))
    return names


def test_archive_round_trip_in_the_reference_format(tmp_path):
    pf = ProbabilityField(10.0, 10.0, 0.0, 0.0)
    pf.expand(-55.0, 42.0, -31.0, 77.0)
    pf.pgrid[3, 4] = 5.0
    pf.total_weight = 7.0
    path = dump_oneka("proj", "t", 0, 10, 365.25, 7, 0.0, (1.0, 2.0), (0.2, 0.3), (20.0, 30.0),
                      [[0.0, 0.0, 0.25, 100.0]], [[1.0, 2.0, 3.0, 0.1]], 50.0, 10.0, 20.0, 1.0,
                      True, 1.0, 10.0, pf, filename=str(tmp_path / "run.bz2"))
    assert ("oneka.probabilityfield", "ProbabilityField") in _globals(path)
    run = load_oneka(path)
    back = run["pfield"]
    assert isinstance(back, ProbabilityField)
    for k in ("deltax", "deltay", "xmin", "xmax", "ymin", "ymax", "nrows", "ncols",
              "total_weight"):
        assert getattr(back, k) == getattr(pf, k)
    assert np.array_equal(back.pgrid, pf.pgrid) and np.array_equal(back.rgrid, pf.rgrid)
    assert run["projectname"] == "proj" and run["maxstep"] == 10.0


def test_archive_loader_refuses_foreign_globals(tmp_path):
    class Evil:
        def __reduce__(self):
            return (os.getcwd, ())
    path = tmp_path / "evil.bz2"
    with bz2.BZ2File(path, "w") as fp:
        pickle.dump({"pfield": Evil()}, fp)
    with pytest.raises(pickle.UnpicklingError):
        load_oneka(str(path))


def test_archive_writer_never_imports_a_module_named_oneka(tmp_path):
    """An importable package named oneka (standing in for the reference) is
    neither imported nor shadowed by dump_oneka, while the archive still names
    oneka.probabilityfield.ProbabilityField, and the plain pickle loader
    (what oneka/archive.py:85-89 uses) resolves it to that package's class."""
    pkg = tmp_path / "site" / "oneka"
    pkg.mkdir(parents=True)
    marker = tmp_path / "imported"
    (pkg / "__init__.py").write_text(f"open({str(marker)!r}, 'a').write('pkg')\n")
    (pkg / "probabilityfield.py").write_text("class ProbabilityField:\n    pass\n")
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = textwrap.dedent(f"""
        import bz2, os, pickle, sys
        sys.path[:0] = [{repo!r}, {str(tmp_path / "site")!r}]
        from onekapy_amd import ProbabilityField
        from onekapy_amd.archive import dump_oneka, load_oneka
        pf = ProbabilityField(10.0, 10.0, 0.0, 0.0)
        pf.expand(-55.0, 42.0, -31.0, 77.0)
        pf.pgrid[2, 3] = 4.0
        pf.total_weight = 4.0
        path = dump_oneka("p", "t", 0, 10, 365.25, 4, 0.0, 1.0, 0.25, 20.0, [], [], 50.0, 10.0,
                          20.0, 0.0, True, 1.0, 10.0, pf, filename={str(tmp_path / "a.bz2")!r})
        assert not any(m == "oneka" or m.startswith("oneka.") for m in sys.modules), sys.modules
        assert not os.path.exists({str(marker)!r})
        back = load_oneka(path)["pfield"]
        assert isinstance(back, ProbabilityField) and back.pgrid[2, 3] == 4.0
        assert not os.path.exists({str(marker)!r})
        with bz2.BZ2File(path) as fp:
            ref = pickle.load(fp)["pfield"]
        import oneka.probabilityfield as opf
        assert type(ref) is opf.ProbabilityField
        assert ref.nrows == pf.nrows and ref.total_weight == 4.0 and ref.pgrid[2, 3] == 4.0
        print("ok")
    """)
    out = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
    assert marker.read_text() == "pkg"   # imported only by the plain loader at the end
    assert ("oneka.probabilityfield", "ProbabilityField") in _globals(str(tmp_path / "a.bz2"))
