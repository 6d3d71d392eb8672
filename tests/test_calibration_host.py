"""The mAs photon-count calibration (chatcarlo/source.py:213-258, tests/test_calibration.py), with
the SpekPy fluences replaced by the test model in heel_model.py (SpekPy is absent): linear in mAs
and field area, the mAs requirement, the SpekPy-absent error, and the heel effect's field-mean
fluence below the central-axis value."""
import math

import numpy as np
import pytest

from heel_model import make_central_fluence, make_provider

BASE = {"kvp": 80.0, "filtration_mm_al": 2.5, "anode_angle_deg": 12.0, "position": [0.0, 0.0, 170.0],
        "direction": [0.0, 0.0, -1.0], "field": {"shape": "rect", "size_cm": [20.0, 20.0], "sid_cm": 100.0}}


@pytest.fixture
def fluence(monkeypatch):
    from chatcarlo_amd import spectrum
    monkeypatch.setattr(spectrum, "fluence_per_mas", make_central_fluence())
    monkeypatch.setattr(spectrum, "heel_spectra_for_source", make_provider())


def test_photon_count_scales_linearly_with_mas(fluence):
    """tests/test_calibration.py:21-26."""
    from chatcarlo_amd.source import photon_count_through_field
    n1 = photon_count_through_field({**BASE, "mas": 1.0})
    n10 = photon_count_through_field({**BASE, "mas": 10.0})
    assert n1 > 0 and n10 == pytest.approx(10.0 * n1, rel=1e-12)


def test_photon_count_scales_with_field_area(fluence):
    """tests/test_calibration.py:29-36, and a cone's area pi r^2."""
    from chatcarlo_amd.source import photon_count_through_field
    small = photon_count_through_field({**BASE, "mas": 1.0, "field": {"size_cm": [10.0, 10.0], "sid_cm": 100.0}})
    large = photon_count_through_field({**BASE, "mas": 1.0, "field": {"size_cm": [20.0, 20.0], "sid_cm": 100.0}})
    assert large == pytest.approx(4.0 * small, rel=1e-12)
    cone = photon_count_through_field({**BASE, "mas": 1.0, "field": {"shape": "cone", "diameter_cm": 20.0,
                                                                      "sid_cm": 100.0}})
    assert cone == pytest.approx(math.pi * 100.0 / 400.0 * large, rel=1e-12)


def test_photon_count_requires_mas(fluence):
    """tests/test_calibration.py:39-42."""
    from chatcarlo_amd.source import photon_count_through_field
    with pytest.raises(ValueError, match="mas"):
        photon_count_through_field(BASE)


def test_photon_count_without_spekpy_raises():
    """chatcarlo/source.py:222-223: no absolute fluence without SpekPy (the Kramers fallback has no
    absolute scale)."""
    try:
        import spekpy  # noqa: F401
        pytest.skip("spekpy is installed")
    except ImportError:
        pass
    from chatcarlo_amd.source import photon_count_through_field
    with pytest.raises(RuntimeError, match="spekpy"):
        photon_count_through_field({**BASE, "mas": 4.0})


def test_heel_field_mean_below_the_central_axis_value(fluence):
    """tests/test_heel_effect.py:113-122: with the heel effect the field-mean fluence is below the
    central-axis value (the anode side loses more than the cathode side gains), and above half of
    it; the mean is that of the heel fluences interpolated over the field."""
    from chatcarlo_amd.source import photon_count_through_field
    src = {**BASE, "kvp": 120.0, "heel_effect": True, "anode_direction": [1.0, 0.0, 0.0],
           "field": {"shape": "rect", "size_cm": [40.0, 40.0], "sid_cm": 100.0}}
    n_heel = photon_count_through_field({**src, "mas": 10.0})
    n_flat = photon_count_through_field({**{k: v for k, v in src.items() if k != "heel_effect"}, "mas": 10.0})
    assert 0.5 * n_flat < n_heel < n_flat
    centres, _, _, flu = make_provider()(src, 100.0, 20.0)
    s = np.linspace(-20.0, 20.0, 101)
    assert n_heel == pytest.approx(np.interp(s, centres, flu).mean() * 10.0 * 1600.0, rel=1e-12)


def test_kvp_spectrum_prefers_spekpy(monkeypatch):
    """chatcarlo/spectrum.py:64-70: a kVp source takes SpekPy's spectrum when SpekPy is there, the
    Kramers fallback with a warning otherwise; a given list wins over both."""
    import warnings
    from chatcarlo_amd import spectrum
    monkeypatch.setattr(spectrum, "spekpy_spectrum", lambda kvp, f, a: (np.array([30.0, 50.0]), np.array([0.25, 0.75])))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        e, cdf = spectrum.spectrum_table(BASE)
    np.testing.assert_array_equal(e, [30.0, 50.0])
    np.testing.assert_allclose(cdf, [0.25, 1.0])
    e2, _ = spectrum.spectrum_table({**BASE, "spectrum": [{"energy_keV": 60.0, "weight": 1.0}]})
    np.testing.assert_array_equal(e2, [60.0])
    monkeypatch.setattr(spectrum, "spekpy_spectrum", lambda kvp, f, a: None)
    with pytest.warns(UserWarning, match="Kramers"):
        e3, _ = spectrum.spectrum_table(BASE)
    assert len(e3) == 60
