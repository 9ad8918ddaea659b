"""Write golden fixtures for the "next" rows by running the OnekaPy reference
(development container only; the reference never travels with this
repository):

    python tests/golden/make_reference_aux.py [REFERENCE_ROOT]

  reference_aux.json      filter_obs cases (utilities.py:34-144) with their
                          outputs; the impact curve's defining values
  reference_impact.npz    a count field and create_impact_plot's (pr, area)
                          (visualize.py:345-390), with the decile rows of
                          oneka.py:284-288
  reference_archive.bz2   an archive written by the reference's dump_oneka
                          (archive.py:59-87)

np.float is aliased to float (numpy 2.x removed it; probabilityfield.py uses
it) and matplotlib draws to the Agg backend.
"""
import glob
import json
import os
import shutil
import sys
import tempfile

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REF)
if not hasattr(np, "float"):
    np.float = float
if not hasattr(np, "bool"):
    np.bool = bool
import matplotlib  # noqa: E402

matplotlib.use("Agg")

from oneka.archive import dump_oneka  # noqa: E402
from oneka.probabilityfield import ProbabilityField  # noqa: E402
from oneka.utilities import filter_obs  # noqa: E402
from oneka.visualize import create_impact_plot  # noqa: E402


def filter_cases():
    wells = [(0.0, 0.0, 0.25, 500.0), (100.0, 0.0, 0.25, 300.0)]
    obs = [
        (10.0, 0.0, 5.0, 0.1),        # within 50 m of well 0: removed
        (50.0, 0.0, 5.1, 0.1),        # exactly 50 m from both wells: removed (<= buffer)
        (200.0, 200.0, 7.0, 0.1),     # run of three within 1 m of the first
        (200.5, 200.0, 7.2, 0.2),
        (200.0, 200.9, 6.9, 0.05),
        (300.0, 300.0, 8.0, 0.3),
        (300.4, 300.0, 8.5, 0.3),     # pair
        (500.0, 500.0, 9.0, 0.1),
        (700.0, 700.0, 9.5, 0.1),
        (500.2, 500.1, 9.3, 0.1),     # near (500, 500) but not consecutive: kept apart
        (800.0, -800.0, 3.0, 0.0),    # zero std, alone: kept as is
    ]
    out = []
    for buffer in (50.0, 10.0):
        out.append({"observations": obs, "wellfield": wells, "buffer": buffer,
                    "retained": [list(map(float, ob)) for ob in filter_obs(obs, wells, buffer)]})
    return out


def impact_case():
    rng = np.random.default_rng(5)
    R = 37
    pf = ProbabilityField(20.0, 20.0, 1000.0, 2000.0)
    pf.expand(1000.0 - 380.0, 1000.0 + 380.0, 2000.0 - 280.0, 2000.0 + 280.0)
    shape = (pf.nrows, pf.ncols)
    counts = rng.integers(0, R + 1, shape).astype(np.float64)
    counts[rng.random(shape) < 0.5] = 0.0
    counts[0, :5] = R
    pf.pgrid = counts
    pf.total_weight = float(R)
    pr, area = create_impact_plot(20.0, pf)
    rows = []
    for p in np.linspace(0.05, 0.95, 19):
        i = np.argmax(pr <= p)
        rows.append((p, pr[i], area[i], area[i] / 4046.86))
    return counts, R, pr, area, np.array(rows), pf


def main():
    aux = {"filter_obs": filter_cases()}
    counts, R, pr, area, rows, pf = impact_case()
    np.savez_compressed(os.path.join(HERE, "reference_impact.npz"), counts=counts,
                        total_weight=float(R), spacing=20.0, pr=pr, area=area, deciles=rows)
    with open(os.path.join(HERE, "reference_aux.json"), "w") as f:
        json.dump(aux, f, indent=1)
    tmp = tempfile.mkdtemp()
    cwd = os.getcwd()
    try:
        os.chdir(tmp)
        os.makedirs("logs", exist_ok=True)
        dump_oneka("golden", "2026-01-01 00:00:00", 0, 100, 3652.5, 37, 0.0, (5.0, 15.0),
                   (0.2, 0.3), (20.0, 30.0), [[1000.0, 2000.0, 0.25, (300.0, 500.0, 700.0)]],
                   [[1100.0, 2100.0, 12.0, 0.1]], 100.0, 20.0, 40.0, 2.0, True, 1.0, 10.0, pf)
        written = glob.glob(os.path.join(tmp, "*.bz2")) + glob.glob(os.path.join(tmp, "logs", "*.bz2"))
        assert len(written) == 1, written
        shutil.copy(written[0], os.path.join(HERE, "reference_archive.bz2"))
    finally:
        os.chdir(cwd)
        shutil.rmtree(tmp)
    print("wrote reference_aux.json, reference_impact.npz, reference_archive.bz2")


if __name__ == "__main__":
    main()
