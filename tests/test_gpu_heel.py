"""The heel effect on the GPU (chatcarlo/source.py:128-149,195-207), with heel spectra from
the test provider in heel_model.py (SpekPy, which makes them in the reference, is absent):
the kept heel coordinates follow the interpolated relative fluence and each heel bin's
energies its spectrum (exact integrals, chi-square), the GPU samples agree with the oracle's
heel sampler, and the reference's tests/test_heel_effect.py checks hold (fewer and harder
photons on the anode side, symmetric without the effect, with cone and rotation, the anode
cut-off).  Batched runs with a heel source equal single-scene runs."""
import numpy as np
import pytest
from scipy import stats

from heel_model import expected_coordinate_probabilities, make_provider

pytestmark = pytest.mark.gpu

SRC = {"kvp": 120.0, "filtration_mm_al": 2.5, "anode_angle_deg": 12.0, "position": [0.0, 0.0, 170.0],
       "direction": [0.0, 0.0, -1.0], "heel_effect": True, "anode_direction": [1.0, 0.0, 0.0],
       "field": {"shape": "rect", "size_cm": [40.0, 40.0], "sid_cm": 100.0}}


@pytest.fixture
def heel(monkeypatch):
    from chatcarlo_amd import spectrum
    prov = make_provider()
    monkeypatch.setattr(spectrum, "heel_spectra_for_source", prov)
    return prov


def _sample(src, n, seed=53):
    from chatcarlo_amd.source import sample_source_photons
    return sample_source_photons(src, n, np.random.default_rng(seed))


def _axis(src):
    from chatcarlo_amd.source import beam_basis, heel_axis_coeffs, heel_span_cm
    d = np.asarray(src["direction"], dtype=float)
    d = d / np.linalg.norm(d)
    u, v = beam_basis(d)
    hu, hv = heel_axis_coeffs(src, d, u, v)
    return d, u, v, hu, hv, heel_span_cm(src["field"], hu, hv)


def _heel_coordinate(src, dirs):
    """A photon's heel coordinate from its direction (static beam): its point on the SID plane."""
    d, u, v, hu, hv, _ = _axis(src)
    t = dirs / (dirs @ d)[:, None] * src["field"]["sid_cm"]
    return (t @ u) * hu + (t @ v) * hv


def test_heel_coordinates_and_energies_match_the_closed_forms_and_the_oracle(heel):
    from oracle.source import sample_source_heel
    n = 400_000
    _, dirs, e = _sample(SRC, n)
    s = _heel_coordinate(SRC, dirs)
    d, u, v, hu, hv, span = _axis(SRC)
    centres, spectra, rel, _ = heel(SRC, 100.0, span)
    edges = np.linspace(-span, span, 41)
    obs = np.histogram(s, edges)[0]
    exp = expected_coordinate_probabilities(edges, centres, rel, -span, span) * n
    chi2 = ((obs - exp) ** 2 / exp).sum()
    print(f"heel coordinates: chi2 {chi2:.1f} over {len(obs)} bins", flush=True)
    assert stats.chi2.sf(chi2, len(obs) - 1) > 1e-3, chi2
    nb = len(centres)
    bins = np.clip(((s + span) / (2 * span / nb)).astype(int), 0, nb - 1)
    grp = np.arange(0.0, 130.0, 10.0)
    for b in range(nb):
        eb, wb = spectra[b]
        got = np.histogram(e[bins == b], grp)[0]
        exp_b = np.histogram(eb, grp, weights=wb)[0] * got.sum()
        keep = exp_b > 5
        chi2 = ((got[keep] - exp_b[keep]) ** 2 / exp_b[keep]).sum()
        assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, (b, chi2)
    # against the oracle's heel sampler (an independent stream): two-sample chi-square of the
    # heel coordinates and of the energies
    _, _, e_o, s_o = sample_source_heel(SRC, n, np.random.default_rng(9), hu, hv, span, centres, spectra, rel)
    for a, b_, ed in ((s, s_o, edges), (e, e_o, grp)):
        table = np.array([np.histogram(a, ed)[0], np.histogram(b_, ed)[0]])
        table = table[:, table.sum(axis=0) > 10]
        assert stats.chi2_contingency(table)[1] > 1e-4


def test_fewer_photons_on_anode_side(heel):
    """tests/test_heel_effect.py:35-41."""
    _, dirs, _ = _sample(SRC, 60_000)
    assert (dirs[:, 0] > 0).sum() < 0.92 * (dirs[:, 0] < 0).sum()


def test_harder_spectrum_on_anode_side(heel):
    """tests/test_heel_effect.py:44-50."""
    _, dirs, e = _sample(SRC, 60_000)
    s = dirs[:, 0] / (-dirs[:, 2]) * 100.0
    assert e[s > 10.0].mean() > e[s < -10.0].mean() + 1.0


def test_heel_off_is_symmetric():
    """tests/test_heel_effect.py:53-57."""
    _, dirs, _ = _sample({**SRC, "heel_effect": False}, 60_000)
    assert abs(int((dirs[:, 0] > 0).sum()) - int((dirs[:, 0] < 0).sum())) < 5 * np.sqrt(len(dirs))


def test_heel_composes_with_cone_and_rotation(heel):
    """tests/test_heel_effect.py:60-74: the anode axis along the rotation axis stays the world x
    axis, so the anode side keeps fewer photons."""
    src = {**SRC, "field": {"shape": "cone", "diameter_cm": 20.0, "sid_cm": 60.0},
           "rotation": {"isocenter": [0.0, 0.0, 110.0], "axis": "x"}}
    _, dirs, e = _sample(src, 20_000)
    assert np.all(np.isfinite(dirs)) and np.all(np.isfinite(e))
    np.testing.assert_allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-5)
    assert (dirs[:, 0] > 0).sum() < (dirs[:, 0] < 0).sum()


def test_anode_cutoff_warns_and_still_samples(monkeypatch):
    """tests/test_heel_effect.py:77-88: bins beyond the cut-off emit nothing; the sampler still
    draws, with no photons there."""
    from chatcarlo_amd import spectrum
    monkeypatch.setattr(spectrum, "heel_spectra_for_source", make_provider())
    src = {**SRC, "field": {"shape": "cone", "diameter_cm": 30.0, "sid_cm": 60.0}}   # half-angle 14 > 12 degrees
    with pytest.warns(UserWarning, match="cut-off"):
        _, dirs, e = _sample(src, 5_000)
    assert np.all(np.isfinite(e)) and np.all(e >= 1.0)
    s = dirs[:, 0] / (-dirs[:, 2] + 1e-12) * 60.0
    assert np.sum(s > 13.0) < 0.01 * len(s)


def test_batched_heel_source_equals_a_single_scene_run(engine, heel):
    """A heel source in a batched run, in its persistent and its event-round form, draws every
    history as a single-scene run does."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.batch import BatchedRun, SceneSpec, Tuning
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    raw = {"source": {**SRC, "position": [0.0, 0.0, 100.0]},
           "geometry": [{"name": "slab", "shape": "box", "material": "water", "center": [0, 0, 0],
                         "size_cm": [30, 30, 10]}], "physics": {"fluorescence": True}}
    grid = (np.array([-20.0, -20.0, -10.0]), (20, 20, 10), 2.0)
    spec = SceneSpec(raw, grid, 30_000, 77)
    engine.set_geometry(Geometry(raw["geometry"]))
    engine.bind_grid(grid[0], grid[1], grid[2], False)
    o = engine.run(SourceRecord(raw["source"]), spec.n_histories, spec.seed, 0, 4096, N.RUN_FLUORESCENCE | N.RUN_GRID)
    k_ref = engine.read_grid(grid[1], moments=False)["kerma"]
    b = BatchedRun([spec], engine)
    for tune in (Tuning(512, tally_copies=2, persistent=True, epoch=2), Tuning(256, 8, 1 << 20)):
        st = b.run(tune)
        assert (st.n_absorbed, st.n_escaped, st.scatter_sum, st.n_events) == (o.n_absorbed, o.n_escaped,
                                                                              o.scatter_sum, o.n_events), tune
        np.testing.assert_allclose(b.read(0)[0], k_ref, rtol=1e-9, atol=1e-12 * k_ref.max())
    # the dose is heavier on the cathode side (x < 0) of the slab
    assert k_ref[:10].sum() > k_ref[10:].sum()


@pytest.mark.parametrize("form,n", [("rounds", 200_000), ("persistent", 200_000), ("rounds", 2_500_000)])
def test_streamed_heel_run_equals_batch_by_batch(engine, heel, form, n):
    """A heel source through the streamed refill (its heel instantiation; 2.5e6 histories run in
    three lanes) and the persistent kernel draws the same photons as the batch-by-batch source
    kernel: equal counters, tallies to FP64 summation order."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    raw = {"source": {**SRC, "position": [0.0, 0.0, 100.0]},
           "geometry": [{"name": "slab", "shape": "box", "material": "water", "center": [0, 0, 0],
                         "size_cm": [30, 30, 10]}]}
    grid = (np.array([-20.0, -20.0, -10.0]), (20, 20, 10), 2.0)
    engine.set_geometry(Geometry(raw["geometry"]))
    rec = SourceRecord(raw["source"])
    res = []
    for extra in ({"rounds": 0, "persistent": N.RUN_PERSISTENT}[form], N.RUN_PER_BATCH):
        engine.bind_grid(grid[0], grid[1], grid[2], True)
        o = engine.run(rec, n, 31, 5_000, 20_000, N.RUN_FLUORESCENCE | N.RUN_GRID | N.RUN_TRACK_UNCERTAINTY | extra)
        g = engine.read_grid(grid[1], moments=True)
        res.append((np.array([o.n_absorbed, o.n_escaped, o.scatter_sum, o.n_fluorescence, o.n_events, o.n_flights,
                              o.n_voxel_visits], dtype=np.int64), g))
    (ca, ga), (cb, gb) = res
    np.testing.assert_array_equal(ca, cb)
    for k in ("kerma", "h10", "kerma_sum2", "h10_sum2"):
        np.testing.assert_allclose(ga[k], gb[k], rtol=1e-9, atol=1e-13 * float(np.abs(gb[k]).max()), err_msg=k)


def test_run_transport_with_mas_calibrates(monkeypatch):
    """tests/test_calibration.py:58-81 with the model fluences in place of SpekPy's: without mas no
    calibration; with mas, n_photons_real is the photon count through the field, with and without
    the heel effect."""
    from heel_model import make_central_fluence
    from chatcarlo_amd import spectrum
    from chatcarlo_amd.scene import validate_scene
    from chatcarlo_amd.source import photon_count_through_field
    from chatcarlo_amd.transport import run_transport
    monkeypatch.setattr(spectrum, "fluence_per_mas", make_central_fluence())
    monkeypatch.setattr(spectrum, "heel_spectra_for_source", make_provider())
    geo = [{"name": "w", "shape": "box", "material": "water", "center": [0, 0, 70], "size_cm": [20, 20, 20]}]
    src = {k: v for k, v in SRC.items() if k not in ("heel_effect", "anode_direction")}
    plain = run_transport(validate_scene({"source": src, "geometry": geo}), n_histories=20_000, seed=2)
    assert plain.n_photons_real is None
    for extra in ({}, {"heel_effect": True, "anode_direction": [1.0, 0.0, 0.0]}):
        s = {**src, **extra, "mas": 4.0}
        sc = validate_scene({"source": s, "geometry": geo})
        assert sc.ok, sc.errors
        r = run_transport(sc, n_histories=20_000, seed=2)
        assert r.n_photons_real == pytest.approx(photon_count_through_field(s), rel=1e-12)
        assert r.n_photons_real > 0
