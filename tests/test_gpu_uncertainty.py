"""The reference's statistical-uncertainty tests through transport (tests/test_uncertainty_transport.py),
on the GPU through run_transport / transport_photons.

Differences kept on purpose: the reference asks totals to be bit-identical with statistics on and
off; here the tallies are FP64 atomic sums, so they agree to summation order (1e-12 relative) and
integer counters exactly.  The reference's worker-count cases are re-expressed as a property of the
(seed, history)-keyed stream: n_workers does not change anything."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RAW = {"source": {"kvp": 100, "position": [0, -50, 0], "direction": [0, 1, 0],
                  "field": {"size_cm": [30, 30], "sid_cm": 100}},
       "geometry": [{"name": "target", "shape": "box", "material": "water", "center": [0, 0, 0],
                     "size_cm": [20, 20, 20]}]}


def _scene():
    from chatcarlo_amd.scene import validate_scene
    sc = validate_scene(RAW)
    assert sc.ok
    return sc


def _run(**kw):
    from chatcarlo_amd.transport import run_transport
    return run_transport(_scene(), **kw)


def test_totals_do_not_depend_on_tracking_uncertainty(engine):
    """test_uncertainty_transport.py:37-66: grid totals, deposits, fractions, scatter and
    fluorescence counts with statistics on and off."""
    on = _run(n_histories=20000, seed=5, dose_grid=True, grid_resolution_cm=5.0, track_uncertainty=True)
    off = _run(n_histories=20000, seed=5, dose_grid=True, grid_resolution_cm=5.0, track_uncertainty=False)
    for a, b in ((on.grid.kerma_keV, off.grid.kerma_keV), (on.grid.h10_track_pSv_cm3, off.grid.h10_track_pSv_cm3)):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300)
    assert on.energy_deposited_MeV.keys() == off.energy_deposited_MeV.keys()
    for k, v in on.energy_deposited_MeV.items():
        assert off.energy_deposited_MeV[k] == pytest.approx(v, rel=1e-12)
    assert (on.fraction_absorbed, on.fraction_escaped, on.mean_scatter_events, on.n_fluorescence,
            on.n_photons_real) == (off.fraction_absorbed, off.fraction_escaped, off.mean_scatter_events,
                                   off.n_fluorescence, off.n_photons_real)


def test_default_tracks_uncertainty(engine):
    """test_uncertainty_transport.py:69-76: the default is track_uncertainty=True."""
    d = _run(n_histories=5000, seed=3, dose_grid=True, grid_resolution_cm=5.0)
    e = _run(n_histories=5000, seed=3, dose_grid=True, grid_resolution_cm=5.0, track_uncertainty=True)
    np.testing.assert_allclose(d.grid.kerma_keV, e.grid.kerma_keV, rtol=1e-12, atol=1e-300)
    assert d.n_batches == e.n_batches
    assert d.energy_deposited_rel_err.keys() == e.energy_deposited_rel_err.keys()


def test_same_seed_reproduces_the_statistics(engine):
    """test_uncertainty_transport.py:79-92, with any worker count: batch moments, hit counts and
    batch counts of two runs with the same seed.  (The reference's two workers make two batches
    of 8000 histories; here n_workers changes nothing, so the batches are set to 2000.)"""
    r1 = _run(n_histories=8000, seed=123, n_workers=2, dose_grid=True, grid_resolution_cm=5.0, batch_size=2000)
    r2 = _run(n_histories=8000, seed=123, n_workers=1, dose_grid=True, grid_resolution_cm=5.0, batch_size=2000)
    np.testing.assert_allclose(r1.grid.kerma_sum2, r2.grid.kerma_sum2, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(r1.grid.h10_sum2, r2.grid.h10_sum2, rtol=1e-12, atol=1e-300)
    np.testing.assert_array_equal(r1.grid.n_batches_hit, r2.grid.n_batches_hit)
    assert r1.n_batches == r2.n_batches == 4
    for k, v in r1.energy_deposited_rel_err.items():
        assert np.isfinite(v)
        assert r2.energy_deposited_rel_err[k] == pytest.approx(v, rel=1e-9)


def _max_kerma_voxel():
    r = _run(n_histories=200_000, seed=999, dose_grid=True, grid_resolution_cm=5.0, batch_size=2000)
    return np.unravel_index(int(np.argmax(r.grid.kerma_keV)), r.grid.kerma_keV.shape)


def test_relative_error_scales_as_one_over_root_n(engine):
    """test_uncertainty_transport.py:105-126: four times the histories halve R (mean of six runs
    at 2e4 and at 8e4 histories, batches of 2000, the reference's bounds 1.3 < ratio < 3)."""
    v = _max_kerma_voxel()

    def mean_r(n, seed0):
        return np.mean([_run(n_histories=n, seed=seed0 + i, dose_grid=True, grid_resolution_cm=5.0,
                             batch_size=2000).grid.kerma_relative_error()[v] for i in range(6)])

    ratio = mean_r(20_000, 300) / mean_r(80_000, 400)
    assert 1.3 < ratio < 3.0, ratio


def test_relative_error_order_does_not_depend_on_batch_size(engine):
    """test_uncertainty_transport.py:129-145: 20 against 10 batches of the same histories."""
    v = _max_kerma_voxel()
    fine = _run(n_histories=40_000, seed=11, dose_grid=True, grid_resolution_cm=5.0, batch_size=2000)
    coarse = _run(n_histories=40_000, seed=12, dose_grid=True, grid_resolution_cm=5.0, batch_size=4000)
    assert fine.n_batches == 20 and coarse.n_batches == 10
    ratio = fine.grid.kerma_relative_error()[v] / coarse.grid.kerma_relative_error()[v]
    assert 0.2 < ratio < 5.0, ratio


def test_deposit_errors_without_grid_and_when_off(engine):
    """test_uncertainty_transport.py:167-197: per-material SEM and R without a dose grid; none,
    and no batches counted, with statistics off; R is NaN with a single batch."""
    r = _run(n_histories=20_000, seed=21, batch_size=2000, dose_grid=False, track_uncertainty=True)
    assert r.grid is None and r.n_batches == 10
    assert r.energy_deposited_rel_err
    for name, rel in r.energy_deposited_rel_err.items():
        assert np.isfinite(rel) and rel >= 0.0
        assert r.energy_deposited_sem_MeV[name] >= 0.0
    off = _run(n_histories=5000, seed=1, dose_grid=False, track_uncertainty=False)
    assert off.energy_deposited_sem_MeV == {} and off.energy_deposited_rel_err == {} and off.n_batches == 0
    one = _run(n_histories=5000, seed=2, batch_size=200_000, dose_grid=False, track_uncertainty=True)
    assert one.n_batches == 1
    assert all(np.isnan(x) for x in one.energy_deposited_rel_err.values())


def test_batch_of_one_history_matches_brute_force(engine):
    """test_uncertainty_transport.py:200-239: one history per batch through transport_photons and
    VoxelGrid.end_batch; the relative error of the hottest voxel equals the brute-force
    std(ddof=1) / sqrt(N) / mean of the per-history contributions."""
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.tally import VoxelGrid
    from chatcarlo_amd.transport import transport_photons
    geo = Geometry([{"name": "slab", "shape": "box", "material": "water", "center": [0.0, 0.0, 0.0],
                     "size_cm": [10.0, 60.0, 60.0]}])
    grid = VoxelGrid.from_bbox(geo.bbox_min, geo.bbox_max, resolution_cm=2.0, track_uncertainty=True)
    rng = np.random.default_rng(17)
    n = 400
    prev = grid.kerma_keV.copy()
    deltas = []
    for _ in range(n):
        transport_photons(np.array([[-10.0, 0.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]), np.array([60.0]), geo,
                          rng, grid=grid)
        grid.end_batch(1)
        deltas.append(grid.kerma_keV - prev)
        prev = grid.kerma_keV.copy()
    deltas = np.stack(deltas)
    v = np.unravel_index(int(np.argmax(grid.kerma_keV)), grid.kerma_keV.shape)
    x = deltas[(slice(None),) + tuple(v)]
    assert x.sum() > 0
    expected = x.std(ddof=1) / np.sqrt(n) / x.mean()
    assert grid.n_batches == n
    assert grid.kerma_relative_error()[v] == pytest.approx(expected, rel=1e-9)
