"""Host-side batch statistics and grid helpers (chatcarlo_amd/tally.py) against the reference's
unit tests of the same functions (tests/test_uncertainty.py:83-200, tests/test_tally.py:96-112)."""
from __future__ import annotations

import numpy as np

from chatcarlo_amd.tally import ScalarMoments, VoxelGrid, combine_moments, relative_error, standard_error


def test_combine_moments_equals_one_shot_aggregation():
    """Moments of two halves of a batch list combine to those of the whole (tests/test_uncertainty.py:83-105)."""
    rng = np.random.default_rng(7)
    sums = rng.gamma(2.0, 3.0, size=37)
    ns = rng.integers(1, 6, size=37)

    def agg(s, n):
        return float(np.sum(s)), float(np.sum(s ** 2 / n)), len(s), int(np.sum(n))

    got = combine_moments(agg(sums[:13], ns[:13]), agg(sums[13:], ns[13:]))
    for g, e in zip(got, agg(sums, ns)):
        assert np.isclose(g, e, rtol=1e-12)
    t = combine_moments((np.array([1.0, 2.0, 3.0]), np.full(3, 0.5), 3, 10), (np.array([4.0, 5.0, 6.0]), np.ones(3), 5, 20))
    assert np.allclose(t[0], [5.0, 7.0, 9.0]) and np.allclose(t[1], 1.5) and t[2:] == (8, 30)


def test_relative_error_undefined_cases_and_clamp():
    """NaN for fewer than two batches and for a zero mean; a variance that rounding makes slightly
    negative is clamped to 0; arrays broadcast voxel by voxel (tests/test_uncertainty.py:122-157)."""
    assert np.isnan(relative_error(T=5.0, Q=25.0, n_batches=1, n_histories=1))
    assert np.isnan(relative_error(T=0.0, Q=0.0, n_batches=5, n_histories=100))
    sem = standard_error(10.0, 10.0 ** 2 / 100 - 1e-12, 4, 100)
    assert np.isfinite(sem) and sem >= 0.0
    r = relative_error(np.array([0.0, 10.0, 20.0, 0.0]), np.array([0.0, 30.0, 90.0, 0.0]), n_batches=5, n_histories=50)
    assert r.shape == (4,) and np.isnan(r[0]) and np.isnan(r[3]) and np.all(np.isfinite(r[[1, 2]]))


def test_scalar_moments():
    """Relative error equals the brute-force SEM / mean at one history per batch; a material
    missing from a batch counts as 0; one batch gives NaN; an empty batch is ignored
    (tests/test_uncertainty.py:158-200)."""
    rng = np.random.default_rng(3)
    per = rng.gamma(2.0, 5.0, size=20)
    sm = ScalarMoments()
    for s in per:
        sm.add_batch({"water": float(s)}, n_histories_in_batch=1000)
    assert 0.0 < sm.relative_errors()["water"] < 1.0
    sm1 = ScalarMoments()
    for s in per:
        sm1.add_batch({"water": float(s)}, n_histories_in_batch=1)
    assert np.isclose(sm1.relative_errors()["water"], per.std(ddof=1) / np.sqrt(len(per)) / per.mean(), rtol=1e-12)
    sm2 = ScalarMoments()
    sm2.add_batch({"water": 10.0, "lead": 5.0}, 100)
    sm2.add_batch({"water": 12.0}, 100)
    sm2.add_batch({"water": 11.0, "lead": 6.0, "aluminum": 1.0}, 100)
    assert sm2.sums["lead"] == 11.0 and sm2.sums["aluminum"] == 1.0
    assert (sm2.n_batches, sm2.n_histories) == (3, 300)
    sm3 = ScalarMoments()
    sm3.add_batch({"water": 42.0}, 1000)
    assert np.isnan(sm3.relative_errors()["water"])
    sm4 = ScalarMoments()
    sm4.add_batch({"water": 42.0}, 0)
    assert (sm4.n_batches, sm4.n_histories) == (0, 0)


def test_grid_end_batch_is_a_no_op_without_tracking_and_h10_map_divides_by_volume():
    """end_batch does nothing unless the grid tracks uncertainty (tests/test_uncertainty.py:214-223);
    the H*(10) map is the track sum over the voxel volume (tests/test_tally.py:96-112)."""
    g = VoxelGrid.from_bbox(np.zeros(3), np.full(3, 10.0), resolution_cm=1.0, track_uncertainty=False)
    g.kerma_keV[0, 0, :3] = 5.0
    g.end_batch(100)
    assert g.kerma_sum2 is None and g.n_batches == 0 and g.n_histories == 0
    g5 = VoxelGrid.from_bbox(np.zeros(3), np.full(3, 5.0), resolution_cm=5.0)
    g5.h10_track_pSv_cm3[0, 0, 0] = 1000 * 0.7 * 3.0
    assert np.isclose(g5.h10_map_pSv().sum(), 1000 * 0.7 * 3.0 / g5.voxel_volume_cm3(), rtol=1e-12)
    assert g5.voxel_volume_cm3() == 125.0
