"""The data pack, h*(10) and the fallback spectrum against the reference's unit tests of the same
quantities (tests/test_materials.py:42-77, tests/test_dose_coefficients.py:17-30,
tests/test_spectrum.py:13-58)."""
from __future__ import annotations

import warnings

import numpy as np
import pytest

from oracle import core


def _muen(pack, name, e):
    m = pack.mat_id(name)
    e = np.atleast_1d(np.asarray(e, dtype=float))
    return core.xs_part(pack, np.full(len(e), m), e, core.MUEN)


def test_water_muen_between_grid_points_and_at_them(pack):
    """mu_en/rho of water falls from 50 through 70 to 80 keV; at the bundled NIST grid points the
    pack returns the table values (tests/test_materials.py:42-56: the reference's own PCHIP passes
    through them to 1e-6; the pack samples that PCHIP on 1024 log-spaced nodes, which leaves 1.1e-5)."""
    from chatcarlo_amd.datapack import load_xaamdi
    v50, v70, v80 = _muen(pack, "water", [50.0, 70.0, 80.0])
    assert v80 < v70 < v50
    e_tab, _, muen_tab = load_xaamdi("water")
    sel = (e_tab >= 1.0) & (e_tab <= 150.0)
    np.testing.assert_allclose(_muen(pack, "water", e_tab[sel]), muen_tab[sel], rtol=2e-5)


def test_water_muen_at_70_kev_is_pchip_not_log_linear(pack):
    """At 70 keV, midway between the 60 and 80 keV grid points, the shape-preserving interpolant
    differs from the log-log chord by more than 0.5 % (tests/test_materials.py:59-70)."""
    ll = np.exp(np.interp(np.log(70.0), np.log([60.0, 80.0]), np.log([3.190e-2, 2.597e-2])))
    assert abs(_muen(pack, "water", 70.0)[0] - ll) / ll > 0.005


def test_water_hvl_at_60_kev(pack):
    """ln 2 / mu of water at 60 keV is within 3.2-3.5 cm (tests/test_materials.py:73-76)."""
    m = pack.mat_id("water")
    mu = core.xs_part(pack, np.array([m]), np.array([60.0]), core.TOTAL)[0] * pack.density[m]
    assert 3.2 < np.log(2.0) / mu < 3.5


def test_h10_shape(pack):
    """h*(10)/Phi rises from 10 to 20 keV, falls towards its minimum near 60 keV and rises again
    by 150 keV (tests/test_dose_coefficients.py:17-24); vector and scalar calls agree."""
    h = core.h10_coeff(pack, np.array([10.0, 20.0, 60.0, 150.0]))
    assert h[1] > h[0] and h[2] < h[1] and h[3] > h[2]
    e = np.array([10.0, 50.0, 100.0])
    assert np.allclose(core.h10_coeff(pack, e), [core.h10_coeff(pack, np.array([x]))[0] for x in e])


def test_fallback_spectrum():
    """The Kramers + aluminium fallback: normalised, non-negative, within (0, kVp], mean energy
    30-60 % of kVp, hardened by more filtration; an explicit spectrum overrides it
    (tests/test_spectrum.py:13-58)."""
    from chatcarlo_amd.spectrum import kramers_fallback_spectrum, spectrum_table
    e, w = kramers_fallback_spectrum(120.0, 2.5)
    assert np.isclose(w.sum(), 1.0) and np.all(w >= 0) and e.min() > 0 and e.max() <= 120.0
    assert 0.3 * 120.0 < np.sum(e * w) < 0.6 * 120.0
    e1, w1 = kramers_fallback_spectrum(100.0, 1.0)
    e5, w5 = kramers_fallback_spectrum(100.0, 5.0)
    assert np.sum(e5 * w5) > np.sum(e1 * w1)
    se, cdf = spectrum_table({"kvp": 80.0, "spectrum": [{"energy_keV": 50.0, "weight": 1.0}]})
    assert se.tolist() == [50.0] and cdf.tolist() == [1.0]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        se, cdf = spectrum_table({"kvp": 80.0, "filtration_mm_al": 2.5})
    assert se.max() <= 80.0 and se.min() > 0 and cdf[-1] == pytest.approx(1.0) and np.all(np.diff(cdf) >= 0)
