"""Impact curve and deciles (GPU count histogram) against the reference's
create_impact_plot values and oneka()'s decile rows."""
import os

import numpy as np
import pytest

from onekapy_amd import ProbabilityField, _lib
from onekapy_amd.impact import CaptureZoneError, count_histogram, decile_table, impact_curve

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _field(counts, total):
    pf = ProbabilityField(20.0, 20.0, 1000.0, 2000.0)
    pf.expand(1000.0 - 380.0, 1000.0 + 380.0, 2000.0 - 280.0, 2000.0 + 280.0)
    pf.pgrid = np.array(counts, dtype=np.float64)
    pf.total_weight = float(total)
    return pf


def test_impact_curve_and_deciles_equal_reference(gpu_ctx):
    ref = np.load(os.path.join(GOLDEN, "reference_impact.npz"))
    pf = _field(ref["counts"], ref["total_weight"])
    pr, area = impact_curve(float(ref["spacing"]), pf)
    assert np.array_equal(pr, ref["pr"])
    assert np.array_equal(area, ref["area"])
    rows = np.array(decile_table(pr, area))
    assert np.array_equal(rows, ref["deciles"])


def test_count_histogram_properties(gpu_ctx):
    rng = np.random.default_rng(3)
    counts = rng.integers(0, 1001, 2_000_003).astype(np.int32)
    counts[rng.random(counts.size) < 0.7] = 0
    hist = count_histogram(counts, 1000)
    assert np.array_equal(hist, np.bincount(counts, minlength=1001))
    assert hist.sum() == counts.size
    with pytest.raises(_lib.OkaError):
        count_histogram(np.array([0, 5, 1001], np.int32), 1000)
    assert np.array_equal(count_histogram(np.zeros(0, np.int32), 3), np.zeros(4, np.int64))


def test_impact_curve_edge_cases(gpu_ctx):
    ref = np.load(os.path.join(GOLDEN, "reference_impact.npz"))
    with pytest.raises(CaptureZoneError):
        impact_curve(20.0, _field(ref["counts"], 0.0))
    with pytest.raises(ValueError):
        impact_curve(20.0, _field(ref["counts"] * 0.5, 37.0))
    # every node certain
    pr, area = impact_curve(20.0, _field(np.full(ref["counts"].shape, 37.0), 37.0))
    assert (pr == 1.0).all() and area[-1] == pr.size * 400.0
