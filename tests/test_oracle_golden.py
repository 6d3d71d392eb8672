"""Pin the oracle and the data pack against the reference's golden values and against mathematics.

Sources of the expected numbers (paths relative to the reference repository):
* NIST XAAMDI spot values: tests/test_materials.py:13-39 (1 %, 2 % for tissues).
* xraylib linear mu logged by the reference's runs: docs/egs5_crosscheck/water20kev/
  run_chatcarlo_water20.log:1 (0.809828), water60kev (0.205873), water150kev (0.150524),
  fluorescence*/chatcarlo_results.json (Pb 100 keV 62.97835651, Pb 150 keV 22.86555,
  Cu 100 keV 4.107923, Cu 20 keV 302.8452, W 100 keV 85.62689 /cm).
* Water 60 keV Compton 0.17703 (EPDL, bound) and 0.18239 (free Klein-Nishina) cm^2/g:
  docs/egs5_crosscheck/check_kn_vs_epdl.log:1-3.
* Bound-Compton mean transfer t_ref = 0.095928 at 60 keV water: tests/test_bound_compton.py:66-80.
* K-shell data: tests/test_fluorescence.py:15-35; Pb K fraction x omega_K = 0.766971 at
  100 keV: docs/plan_fluorescence.md:209-212.
* ICRP-74 table points: tests/test_dose_coefficients.py:9-14.
The partial-cross-section split, F(Z,q), S(Z,q) and compositions come from the
pack's own atomic model (xraylib is unavailable) and are pinned only through
these anchors and the closed-form checks below.
"""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import core
from oracle.philox import philox4x32_10

RNG = np.random.default_rng(7)


def mu_lin(pack, name, e):
    m = pack.mat_id(name)
    return float(core.xs_lookup(pack, np.array([m]), np.array([float(e)]))["mu"][0])


@pytest.mark.parametrize("mat,e,ref_mu,ref_muen", [
    ("water", 60.0, 2.059e-1, 3.190e-2), ("water", 100.0, 1.707e-1, 2.546e-2),
    ("aluminum", 60.0, 2.778e-1, 1.099e-1), ("aluminum", 100.0, 1.704e-1, 3.794e-2),
    ("lead", 100.0, 5.549e0, 1.976e0), ("soft_tissue", 60.0, 2.048e-1, 3.264e-2),
    ("bone", 60.0, 3.148e-1, 1.400e-1), ("air", 60.0, 1.875e-1, 3.041e-2)])
def test_nist_spot_values(pack, mat, e, ref_mu, ref_muen):
    m = pack.mat_id(mat)
    r = core.xs_lookup(pack, np.array([m]), np.array([e]))
    rho = pack.density[m]
    rel = 0.02 if mat in ("soft_tissue", "bone", "lung") else 0.01
    assert r["mu"][0] / rho == pytest.approx(ref_mu, rel=rel)
    assert r["mu_en"][0] / rho == pytest.approx(ref_muen, rel=1e-3)


@pytest.mark.parametrize("mat,e,golden", [
    ("water", 20.0, 0.809828), ("water", 60.0, 0.205873), ("water", 150.0, 0.150524),
    ("lead", 100.0, 62.97835651), ("lead", 150.0, 22.86555), ("copper", 100.0, 4.107923),
    ("copper", 20.0, 302.8452), ("tungsten", 100.0, 85.62689)])
def test_reference_logged_linear_mu(pack, mat, e, golden):
    assert mu_lin(pack, mat, e) == pytest.approx(golden, rel=0.01)


def test_water_60kev_compton_bound_and_free(pack):
    from chatcarlo_amd import atomic as A
    from chatcarlo_amd.datapack import element_scatter_xs, energy_nodes
    e = energy_nodes()
    j = int(np.argmin(np.abs(e - 60.0)))
    w = dict(A.MATERIALS["water"][1])
    bound = sum(w[z] * element_scatter_xs(z)[1][j] for z in w)
    assert bound == pytest.approx(0.17703 * (e[j] / 60.0) ** -0.12, rel=0.015)
    # free-electron Klein-Nishina per gram: closed form sigma_KN * Z N_A / A
    a = 60.0 / 510.99895
    re2 = A.R_E_CM ** 2
    kn = 2 * math.pi * re2 * ((1 + a) / a ** 2 * (2 * (1 + a) / (1 + 2 * a) - math.log(1 + 2 * a) / a)
                              + math.log(1 + 2 * a) / (2 * a) - (1 + 3 * a) / (1 + 2 * a) ** 2)
    free = sum(w[z] * kn * z * A.N_AVOGADRO / A.ATOMIC_WEIGHT[z] for z in w)
    assert free == pytest.approx(0.18239, rel=2e-3)


def test_water_60kev_split_against_the_reference_epdl_values(pack):
    """The reference's xraylib (EPDL) split of water at 60 keV: photo 0.014926, Compton 0.177029,
    Rayleigh 0.013918, total 0.205873 cm^2/g (docs/egs5_crosscheck/plan_compton_transfer_check.md:
    14-19).  The residual photo alone (NIST total - model scatter) is 5.4 % low and fails; the
    mu_en/rho-based estimate with the scatter closing the total meets it."""
    m = pack.mat_id("water")
    r = core.xs_lookup(pack, np.array([m]), np.array([60.0]))
    assert r["photo"][0] == pytest.approx(0.014926, rel=5e-3)
    assert r["compt"][0] == pytest.approx(0.177029, rel=1e-3)
    assert r["rayl"][0] == pytest.approx(0.013918, rel=5e-3)
    assert r["mu"][0] == pytest.approx(0.205873, rel=5e-4)


def test_compton_transfer_matches_the_reference_quadrature(pack):
    """sigma_tr / sigma_incoh of water at 60 keV is the S(q)-weighted mean transfer t_ref = 0.095928
    of docs/egs5_crosscheck/check_compton_transfer.py:45-84 (the reference's quadrature with
    xraylib S(q)); the pack's Slater S(q) gives it 0.3 % low (0.77 = d ln t / d ln E there)."""
    from chatcarlo_amd import atomic as A
    from chatcarlo_amd.datapack import element_compton_transfer, element_scatter_xs, energy_nodes
    e = energy_nodes()
    j = int(np.argmin(np.abs(e - 60.0)))
    w = dict(A.MATERIALS["water"][1])
    tr = sum(w[z] * element_compton_transfer(z)[j] for z in w)
    inc = sum(w[z] * element_scatter_xs(z)[1][j] for z in w)
    assert tr / inc == pytest.approx(0.095928 * (e[j] / 60.0) ** 0.77, rel=5e-3)


def test_h10_table_points(pack):
    known = {10.0: 0.061, 30.0: 0.81, 60.0: 0.51, 100.0: 0.61, 150.0: 0.89}
    for e, v in known.items():
        assert core.h10_coeff(pack, e)[0] == pytest.approx(v, rel=1e-6)
        assert core.h10_coeff_dev(pack, e)[0] == pytest.approx(v, rel=1e-6)
    e = np.exp(RNG.uniform(np.log(10), np.log(150), 1000))
    np.testing.assert_allclose(core.h10_coeff_dev(pack, e), core.h10_coeff(pack, e), rtol=1e-12)


def test_k_shell_data(pack):
    pb = pack.elements.index(82)
    w = pack.elements.index(74)
    assert abs(pack.fl_edge[pb] - 88.00) < 0.1
    assert abs(pack.fl_omega[pb] - 0.96) < 0.02
    probs = np.diff(np.concatenate([[0.0], pack.fl_cdf[pb]]))
    assert abs(pack.fl_lines[pb][np.argmax(probs)] - 74.97) < 0.1
    assert abs(pack.fl_edge[w] - 69.53) < 0.1
    probs_w = np.diff(np.concatenate([[0.0], pack.fl_cdf[w]]))
    assert abs(pack.fl_lines[w][np.argmax(probs_w)] - 59.32) < 0.1
    for z in (26, 29, 74, 82):
        i = pack.elements.index(z)
        assert pack.fl_cdf[i][-1] == 1.0
    # K-shell share x omega_K for Pb at 100 keV: the reference's CS_Photo_Partial / CS_Photo
    # gives 0.796109 x 0.9634 = 0.766971 (docs/plan_fluorescence.md:209-212).  The pack's
    # edge-jump continuation of the NIST photo cross section gives 0.79297: the 0.4 % gap is the
    # K-edge jump of NIST XCOM (4.75) against EPDL97's; 0.5 % is that gap with its rounding.
    assert core.k_share(pack, pb, 100.0) == pytest.approx(0.796109, rel=5e-3)
    assert core.k_share(pack, pb, 100.0) * pack.fl_omega[pb] == pytest.approx(0.766971, rel=5e-3)


@pytest.mark.parametrize("z,mat,share,omega,photo", [(29, "copper", 0.896, 0.454, 0.628),
                                                     (74, "tungsten", 0.816, 0.954, 0.936)])
def test_k_shell_share_yield_and_photo_fraction_at_100kev(pack, z, mat, share, omega, photo):
    """xraylib's K-shell share, omega_K and photoelectric fraction at 100 keV as the reference logs
    them (docs/egs5_crosscheck/fluorescence_copper/PREREGISTRATION.md:56,
    fluorescence_tungsten/PREREGISTRATION.md:39-40, three digits each).  A share held constant at
    its edge value (Cu 0.867) or continued as a straight power law (Cu 0.854) fails the copper
    case; the Krause yields (Cu 0.440, W 0.950) fail the yields."""
    i = pack.elements.index(z)
    assert core.k_share(pack, i, 100.0) == pytest.approx(share, abs=1e-3)
    assert pack.fl_omega[i] == pytest.approx(omega, abs=1e-3)
    m = pack.mat_id(mat)
    r = core.xs_lookup(pack, np.array([m]), np.array([100.0]))
    assert r["p_photo"][0] == pytest.approx(photo, rel=1e-2)


def test_bound_compton_mean_transfer(pack):
    n, k = 200_000, 40
    m = pack.mat_id("water")
    u = RNG.random((n, 1 + 3 * k))
    eps, cos, el, tr = core.compton_u(pack, np.full(n, m), np.full(n, 60.0), u, k)
    assert np.all(tr > 0)
    t = 1.0 - eps
    se = t.std() / math.sqrt(n)
    # reference EPDL anchor; the Slater-orbital S(Z,q) is allowed 0.5 % on top of 3 sigma
    assert abs(t.mean() - 0.095928) < 3 * se + 0.005 * 0.095928


def test_compton_free_limit_is_klein_nishina(pack):
    """With S/Z = 1 the bound sampler reduces to Kahn's Klein-Nishina sampler: KS vs the closed-form CDF."""
    p = pack.rounded()
    p.sq_over_z = np.ones_like(p.sq_over_z)
    n, k = 40_000, 40
    E = 60.0
    u = RNG.random((n, 1 + 3 * k))
    eps, cos, _, tr = core.compton_u(p, np.full(n, p.mat_id("water")), np.full(n, E), u, k)
    a = E / 511.0
    grid = np.linspace(-1, 1, 4001)
    pp = 1.0 / (1.0 + a * (1 - grid))
    pdf = pp * pp * (pp + 1 / pp - (1 - grid ** 2))
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    res = stats.kstest(cos, lambda x: np.interp(x, grid, cdf))
    assert res.pvalue > 1e-3
    # eps and cos obey Compton kinematics exactly
    np.testing.assert_allclose(eps, 1.0 / (1.0 + a * (1.0 - cos)), rtol=1e-12)


def test_rayleigh_matches_form_factor_distribution(pack):
    """Sampled cos theta vs the CDF of (1 + c^2) F(Z,q)^2 by quadrature (oxygen in water, 30 keV)."""
    from chatcarlo_amd import atomic as A
    n, k = 40_000, 32
    E = 30.0
    m = pack.mat_id("water")
    u = RNG.random((n, 1 + 2 * k))
    cos, el, tr = core.rayleigh_u(pack, np.full(n, m), np.full(n, E), u, k)
    oxy = pack.elements.index(8)
    c = cos[el == oxy]
    grid = np.linspace(-1, 1, 20001)
    q = E / A.HC_KEV_ANGSTROM * np.sqrt((1 - grid) / 2)
    f = np.interp(q, np.linspace(0, 20, 2000), A.form_factor(8, np.linspace(0, 20, 2000)))
    pdf = (1 + grid ** 2) * f ** 2
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    res = stats.kstest(c, lambda x: np.interp(x, grid, cdf))
    assert res.pvalue > 1e-3
    assert np.all(tr > 0) and np.mean(tr) < 2.0   # acceptance >= 50 %


def test_fluorescence_rate_lead_100kev(pack):
    n = 100_000
    m = pack.mat_id("lead")
    u = RNG.random((n, 4))
    el, emit, eline = core.fluor_u(pack, np.full(n, m), np.full(n, 100.0), u)
    pb = pack.elements.index(82)
    p = float(core.k_share(pack, pb, 100.0)) * pack.fl_omega[pb]
    assert abs(emit.mean() - p) < 4 * math.sqrt(p * (1 - p) / n)
    assert np.all(eline[emit] >= 5.0) and np.all(eline[emit] < 100.0)
    el2, emit2, _ = core.fluor_u(pack, np.full(1000, m), np.full(1000, 87.0), RNG.random((1000, 4)))
    assert not emit2.any()


def test_rotation_and_isotropy():
    n = 50_000
    d = RNG.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:100] = [0, 0, 1]
    c = RNG.uniform(-1, 1, n)
    nd = core.rotate_u(d, c, RNG.random(n))
    np.testing.assert_allclose(np.linalg.norm(nd, axis=1), 1.0, rtol=1e-12)
    np.testing.assert_allclose(np.sum(nd * d, axis=1), c, atol=1e-12)
    iso = core.isotropic_u(RNG.random(n), RNG.random(n))
    assert abs(iso.mean(axis=0)).max() < 0.02
    assert stats.kstest(iso[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-3


def test_philox_known_answers():
    c = np.array([[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]],
                 dtype=np.uint32)
    k = np.array([[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]], dtype=np.uint32)
    want = np.array([[0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8],
                     [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD],
                     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]], dtype=np.uint32)
    np.testing.assert_array_equal(philox4x32_10(c, k), want)


def test_xs_lookup_is_loglog_linear_between_nodes(pack):
    """Away from edges the lookup equals np.interp of ln sigma over ln E (chatcarlo/materials.py:240-253)."""
    e = np.exp(RNG.uniform(0, np.log(150), 5000))
    for name in ("water", "bone", "lead", "concrete"):
        m = pack.mat_id(name)
        i, _ = core.energy_interval(pack, e)
        plain = pack.edge_of_interval[m, i] < 0
        for part in (core.PHOTO, core.COMPT, core.RAYL, core.MUEN):
            got = core.xs_part(pack, np.full(len(e), m), e, part)
            ref = np.exp(np.interp(np.log(e), np.log(pack.e_node), np.log(pack.sig[m, part])))
            np.testing.assert_allclose(got[plain], ref[plain], rtol=1e-10)
        # at nodes the lookup returns the node values exactly
        got = core.xs_part(pack, np.full(len(pack.e_node), m), pack.e_node, core.COMPT)
        np.testing.assert_allclose(got, pack.sig[m, core.COMPT], rtol=1e-12)


def test_edge_jumps_are_kept(pack):
    m = pack.mat_id("lead")
    ee = 88.0045
    below = core.xs_part(pack, np.array([m]), np.array([ee * (1 - 1e-7)]), core.PHOTO)[0]
    above = core.xs_part(pack, np.array([m]), np.array([ee * (1 + 1e-5)]), core.PHOTO)[0]
    assert above / below > 3.5


def test_lead_k_edge_jump_is_sharp(pack):
    """Photoelectric jump across lead's K edge, 87.9 against 88.1 keV, above 4 as the reference's
    tests/test_xs_tables.py:99-112 requires of its tables (the pack gives 4.7)."""
    m = pack.mat_id("lead")
    below, above = core.xs_part(pack, np.array([m, m]), np.array([87.9, 88.1]), core.PHOTO)
    assert above / below > 4.0


def test_energy_range_error(pack):
    with pytest.raises(ValueError):
        core.xs_lookup(pack, np.array([0]), np.array([0.5]))
    with pytest.raises(ValueError):
        core.xs_lookup(pack, np.array([0]), np.array([151.0]))


# ---- the reference's Rayleigh invariants (tests/test_rayleigh.py), on the pack and the oracle sampler

def test_form_factor_at_zero_momentum_is_z():
    """F(Z, 0) = Z: every electron scatters in phase (reference tests/test_rayleigh.py:20-23)."""
    from chatcarlo_amd import atomic as A
    for z in (1, 8, 20, 82):
        assert abs(float(A.form_factor(z, np.array([0.0]))[0]) - z) < 1e-6


def test_form_factor_decreases_with_momentum(pack):
    """F(82, q) never rises over the pack's q grid (reference tests/test_rayleigh.py:26-28)."""
    from chatcarlo_amd import atomic as A
    q = np.sqrt(pack.ray_x)
    assert np.all(np.diff(A.form_factor(82, q)) <= 1e-9)


def test_rayleigh_element_selection_in_bone(pack):
    """Bone's Rayleigh element-selection CDF picks only bone's elements, ends at 1 at 30, 60 and
    100 keV, and picks calcium more often than its mass fraction, the coherent cross section
    rising steeply with Z (reference tests/test_rayleigh.py:65-78,106-116)."""
    m = pack.mat_id("bone")
    ca = list(pack.mat_elem[m]).index(pack.elements.index(20))
    for e in (30.0, 50.0, 60.0, 100.0):
        i, f = core.energy_interval(pack, np.array([e]))
        cdf = core.sel_cdf_at(pack, m, 2, i[0], f[0])
        assert abs(cdf[-1] - 1.0) < 1e-12 and np.all(np.diff(cdf) >= 0)
        p_ca = cdf[ca] - (cdf[ca - 1] if ca > 0 else 0.0)
        assert p_ca > pack.mat_frac[m][ca], (e, p_ca, pack.mat_frac[m][ca])
    n = 5000
    i, f = core.energy_interval(pack, np.full(n, 50.0))
    picked = core.select_elements(pack, np.full(n, m), 2, i, f, RNG.random(n))
    assert set(np.unique(picked).tolist()) <= set(pack.mat_elem[m])


# ---- the reference's bound-Compton and fluorescence invariants (tests/test_bound_compton.py,
#      tests/test_fluorescence.py) on the pack

def test_incoherent_function_runs_from_zero_to_z(pack):
    """S(Z, 0) = 0, S(Z, q) rises monotonically, and S/Z is within 1 % of 1 at the table's end
    (reference tests/test_bound_compton.py:17-26)."""
    for z in (1, 8, 20, 82):
        s = pack.sq_over_z[pack.elements.index(z)]
        assert s[0] == 0.0
        assert abs(s[-1] - 1.0) < 0.01
        assert np.all(np.diff(s) >= -1e-9)


def _share(pack, name, kind, z, e):
    m = pack.mat_id(name)
    i, f = core.energy_interval(pack, np.array([e]))
    cdf = core.sel_cdf_at(pack, m, kind, i[0], f[0])
    k = list(pack.mat_elem[m]).index(pack.elements.index(z))
    assert abs(cdf[-1] - 1.0) < 1e-12 and np.all(np.diff(cdf) >= 0)
    return cdf[k] - (cdf[k - 1] if k > 0 else 0.0), pack.mat_frac[m][k]


def test_compton_picks_hydrogen_above_its_mass_fraction_in_bone(pack):
    """Incoherent scattering per gram follows Z/A, largest for hydrogen, so bone's Compton element
    selection picks H more often than its mass fraction (reference tests/test_bound_compton.py:99-112)."""
    p, frac = _share(pack, "bone", 1, 1, 50.0)
    assert p > frac


def test_photo_selection_in_water_is_oxygen(pack):
    """Water's photoelectric element selection is > 90 % oxygen at 50 keV (reference
    tests/test_fluorescence.py:38-42)."""
    p, _ = _share(pack, "water", 0, 8, 50.0)
    assert p > 0.9


def test_k_line_probabilities_sum_to_one(pack):
    """Line probabilities of every fluorescing element sum to 1 (reference tests/test_fluorescence.py:31-35)."""
    for z in (20, 26, 29, 74, 82):
        i = pack.elements.index(z)
        if pack.fl_edge[i] > 0:
            probs = np.diff(np.concatenate([[0.0], pack.fl_cdf[i]]))
            assert abs(probs.sum() - 1.0) < 1e-12 and np.all(probs >= -1e-15)
