"""Impact curve and deciles (GPU count histogram, or GPU radix sort for
fields of other weights) against the reference's create_impact_plot values
and oneka()'s decile rows, and the smoothed probability map against
scipy.ndimage.gaussian_filter (visualize.py:232-233)."""
import os

import numpy as np
import pytest

from onekapy_amd import ProbabilityField, _lib
from onekapy_amd.impact import (CaptureZoneError, count_histogram, decile_table, impact_curve,
                                smoothed_probability, sorted_probabilities)

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
    # a field of weight-0.5 realizations is sorted as a whole (no histogram)
    pr, _ = impact_curve(20.0, _field(ref["counts"] * 0.5, 37.0))
    assert np.array_equal(pr, np.flip(np.sort(ref["counts"] * 0.5 / 37.0, axis=None)))
    # every node certain
    pr, area = impact_curve(20.0, _field(np.full(ref["counts"].shape, 37.0), 37.0))
    assert (pr == 1.0).all() and area[-1] == pr.size * 400.0


def test_impact_curve_of_weighted_fields_equals_numpy_sort(gpu_ctx):
    """Fields that do not hold whole counts: the reference's own two-call
    capture zone (weights 1 and 0.5, tests/golden/reference_vectors.npz) and
    random fractional fields up to 2000 x 2000 nodes give exactly the array
    visualize.py:382 builds with numpy's sort."""
    ref = np.load(os.path.join(GOLDEN, "reference_vectors.npz"))
    pg = ref["cz_pgrid"]
    pf = _field(np.zeros((29, 39)), 1.0)
    pf.pgrid, pf.total_weight = pg.astype(np.float64), 1.5
    assert not np.array_equal(pg, np.rint(pg))
    pr, area = impact_curve(10.0, pf)
    assert np.array_equal(pr, np.flip(np.sort(pg / 1.5, axis=None)))
    assert np.array_equal(area, (np.arange(pg.size) + 1) * 100.0)
    rng = np.random.default_rng(7)
    for shape, tw in (((1, 1), 3.0), ((257, 311), 0.75), ((2000, 2000), 1234.5)):
        g = rng.random(shape) * tw
        g[rng.random(shape) < 0.6] = 0.0
        got = sorted_probabilities(g, tw)
        assert np.array_equal(got, np.flip(np.sort(g / tw, axis=None))), shape


@pytest.mark.parametrize("smooth", [0.0, 0.5, 1.0, 2.0, 3.7, 8.0])
def test_smoothed_probability_equals_scipy(gpu_ctx, smooth):
    """Pr = pgrid / total_weight and its Gaussian smoothing (sigma in nodes,
    zeros beyond the lattice) equal scipy.ndimage.gaussian_filter(Z, smooth,
    mode='constant', cval=0.0) bit for bit, on the reference's impact field
    and on a random 600 x 713 field."""
    from scipy.ndimage import gaussian_filter
    ref = np.load(os.path.join(GOLDEN, "reference_impact.npz"))
    rng = np.random.default_rng(int(smooth * 10))
    big = rng.integers(0, 60, (600, 713)).astype(np.float64)
    big[rng.random(big.shape) < 0.5] = 0.0
    for counts, total in ((ref["counts"], float(ref["total_weight"])), (big, 59.0)):
        pf = _field(counts, total)
        z = counts / total
        want = gaussian_filter(z, smooth, mode="constant", cval=0.0) if smooth > 0 else z
        got = smoothed_probability(pf, smooth)
        assert np.array_equal(got, want)
    with pytest.raises(CaptureZoneError):
        smoothed_probability(_field(ref["counts"], 0.0), smooth)
