"""The heel effect on the host: the oracle's heel sampler pinned by closed forms, the sampler
tables built from heel spectra, and the behaviour without SpekPy (chatcarlo/source.py:43-65,
128-149, 195-207; spectrum.py:88-152)."""
import numpy as np
import pytest
from scipy import stats

from heel_model import expected_coordinate_probabilities, make_provider

SRC = {"kvp": 120.0, "filtration_mm_al": 2.5, "anode_angle_deg": 12.0, "position": [0.0, 0.0, 170.0],
       "direction": [0.0, 0.0, -1.0], "heel_effect": True, "anode_direction": [1.0, 0.0, 0.0],
       "field": {"shape": "rect", "size_cm": [40.0, 40.0], "sid_cm": 100.0}}


def _axis(src):
    from chatcarlo_amd.source import beam_basis, heel_axis_coeffs, heel_span_cm
    d = np.asarray(src["direction"], dtype=float)
    d = d / np.linalg.norm(d)
    u, v = beam_basis(d)
    hu, hv = heel_axis_coeffs(src, d, u, v)
    return hu, hv, heel_span_cm(src["field"], hu, hv)


def test_heel_axis_and_span():
    """chatcarlo/source.py:43-64: the anode direction +x across a beam along -z is +-u of the
    beam basis, so the span of a 40 cm field is 20 cm; a tilted anode direction loses its
    component along the beam; a cone's span is its radius on the SID plane."""
    from chatcarlo_amd.source import heel_span_cm
    hu, hv, span = _axis(SRC)
    assert hu ** 2 + hv ** 2 == pytest.approx(1.0)
    assert abs(hu) == pytest.approx(1.0) and span == pytest.approx(20.0)
    hu2, hv2, _ = _axis({**SRC, "anode_direction": [1.0, 0.0, 5.0]})
    assert (hu2, hv2) == pytest.approx((hu, hv))
    hu3, hv3, span3 = _axis({**SRC, "anode_direction": [1.0, 1.0, 0.0], "field": {"size_cm": [30.0, 10.0],
                                                                                 "sid_cm": 100.0}})
    assert span3 == pytest.approx(abs(hu3) * 15.0 + abs(hv3) * 5.0)
    assert heel_span_cm({"shape": "cone", "diameter_cm": 20.0, "sid_cm": 60.0}, hu, hv) == 10.0
    with pytest.raises(ValueError, match="parallel"):
        _axis({**SRC, "anode_direction": [0.0, 0.0, 1.0]})


def test_oracle_heel_coordinates_follow_the_relative_fluence():
    """With the heel axis along one side of a rect field, the heel coordinate of a drawn point is
    uniform on [-span, span]; the kept ones have the density of the interpolated relative
    fluence (an exact integral, chi-square over 40 bins)."""
    from oracle.source import sample_source_heel
    hu, hv, span = _axis(SRC)
    centres, spectra, rel, _ = make_provider()(SRC, 100.0, span)
    rng = np.random.default_rng(3)
    _, _, _, s = sample_source_heel(SRC, 200_000, rng, hu, hv, span, centres, spectra, rel)
    edges = np.linspace(-span, span, 41)
    obs = np.histogram(s, edges)[0]
    exp = expected_coordinate_probabilities(edges, centres, rel, -span, span) * len(s)
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert stats.chi2.sf(chi2, len(obs) - 1) > 1e-3, chi2
    # the anode side (s > 0) is the thinner one
    assert (s > 0).sum() < 0.9 * (s < 0).sum()


def test_oracle_heel_energies_follow_their_bin_spectrum():
    """Each photon's energy comes from the spectrum of its heel bin: chi-square per bin against
    that bin's weights (over 10 keV groups), and the anode-side bins are harder."""
    from oracle.source import sample_source_heel
    hu, hv, span = _axis(SRC)
    centres, spectra, rel, _ = make_provider()(SRC, 100.0, span)
    rng = np.random.default_rng(4)
    _, _, e, s = sample_source_heel(SRC, 300_000, rng, hu, hv, span, centres, spectra, rel)
    nb = len(centres)
    bins = np.clip(((s + span) / (2 * span / nb)).astype(int), 0, nb - 1)
    means = []
    for b in (0, nb // 2, nb - 1):
        eb, wb = spectra[b]
        grp = np.arange(0.0, 130.0, 10.0)
        obs = np.histogram(e[bins == b], grp)[0]
        exp = np.histogram(eb, grp, weights=wb)[0] * obs.sum()
        keep = exp > 5
        chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
        assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3, (b, chi2)
        means.append(e[bins == b].mean())
    assert means[0] < means[1] < means[2]


def test_heel_tables_from_spectra():
    """The sampler tables: one energy grid over the emitting energies, each bin's CDF on it
    (drawing from it gives the bin's weights), bins beyond the cut-off take the nearest emitting
    bin's CDF and are outside [lo, hi]."""
    from chatcarlo_amd.source import heel_tables
    src = {**SRC, "field": {"shape": "cone", "diameter_cm": 30.0, "sid_cm": 60.0}}
    with pytest.warns(UserWarning, match="cut-off"):   # atan(15 / 60) > 12 degrees
        centres, spectra, rel, _ = make_provider()(src, 60.0, 15.0)
    grid, cdf, rel2, lo, hi = heel_tables(centres, spectra, rel)
    assert grid[0] >= 1.0 and np.all(np.diff(grid) > 0)
    assert lo == 0 and hi == int(np.flatnonzero(rel > 0)[-1]) and hi < len(centres) - 1
    np.testing.assert_array_equal(rel2, rel)
    for k in range(len(centres)):
        assert np.all(np.diff(cdf[k]) >= 0) and cdf[k, -1] == pytest.approx(1.0)
    e3, w3 = spectra[3]
    on = w3 > 0
    np.testing.assert_allclose(np.diff(np.concatenate([[0.0], cdf[3]])), w3[on] / w3[on].sum(), atol=1e-12)
    np.testing.assert_array_equal(cdf[hi + 1], cdf[hi])


def test_heel_without_spekpy_raises_as_the_reference():
    """chatcarlo/spectrum.py:100-102: without spekpy the heel spectra cannot be made."""
    pytest.importorskip("numpy")
    try:
        import spekpy  # noqa: F401
        pytest.skip("spekpy is installed")
    except ImportError:
        pass
    from chatcarlo_amd.source import SourceRecord
    with pytest.raises(RuntimeError, match="spekpy"):
        SourceRecord(SRC)


def test_heel_source_record(monkeypatch):
    """A heel source fills the cc_source heel fields from the provider's spectra."""
    from chatcarlo_amd import spectrum
    from chatcarlo_amd.source import SourceRecord
    monkeypatch.setattr(spectrum, "heel_spectra_for_source", make_provider())
    rec = SourceRecord(SRC)
    s = rec.struct
    assert s.heel_bins == 15 and (s.heel_nz_lo, s.heel_nz_hi) == (0, 14)
    assert s.heel_span == pytest.approx(20.0) and abs(s.heel_hu) == pytest.approx(1.0)
    assert rec.heel_cdf.shape == (15, s.n_spec) and rec.heel_rel_flu.max() == 1.0
    plain = SourceRecord({k: v for k, v in SRC.items() if k != "heel_effect"})
    assert plain.struct.heel_bins == 0
