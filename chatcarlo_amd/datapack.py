"""Physics data pack: every table the transport path reads, built on the host.

The reference tabulates per-element photo/Compton/Rayleigh cross sections on a
log-energy grid with edge bracketing points and interpolates them log-log
linearly (chatcarlo/materials.py:161-262), mixes them by mass fraction
(:265-275), draws the interacting element by mass fraction x element cross
section (:358-440), reads mu_en/rho from the bundled NIST XAAMDI tables with a
log-log PCHIP (:299-335), and samples Rayleigh and bound Compton from
F(Z,q)/S(Z,q) tables on q in [0, 20] 1/Angstrom (:375-406, :483-496).

This module rebuilds the same set of tables for one GPU-friendly layout:

* one energy grid for every material: N_E nodes uniform in ln E over
  [1, 150] keV, so the interval index is arithmetic, plus per-material edge
  records (edge energy, two-sided values) for intervals that hold an
  absorption edge, which keeps a jump a jump exactly as the reference's
  +-1e-5 bracketing points do;
* per material and node: sigma_photo, sigma_compton, sigma_rayleigh [cm^2/g]
  and mu_en/rho, with the log slope to the next node; the lookup is log-log
  linear per part, as in chatcarlo/materials.py:253;
* per material, interaction type and node: cumulative element-selection
  fractions;
* per element: S(Z,q)/Z on the reference's q grid, Rayleigh cumulative
  A(x) = int_0^x F^2 dx' on x = q^2, K-shell fluorescence data.

Data sources (xraylib is unavailable offline): total mu/rho and mu_en/rho are
the NIST XAAMDI values bundled with the reference; coherent and incoherent
cross sections are integrals of the Slater-orbital F and S of
chatcarlo_amd.atomic over the Thomson and Klein-Nishina distributions;
photoelectric = NIST total - coherent - incoherent, interpolated log-log
between NIST nodes within edge-free segments.
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np
from scipy.interpolate import PchipInterpolator

from . import atomic as A

DATA_DIR = Path(__file__).resolve().parent / "physdata"

E_MIN_KEV = 1.0      # chatcarlo/materials.py:95
E_MAX_KEV = 150.0    # chatcarlo/materials.py:96
N_E = 1024
LN_E0 = math.log(E_MIN_KEV)
DLN_E = (math.log(E_MAX_KEV) - LN_E0) / (N_E - 1)
MAX_EDGES = 16

# Rayleigh / Compton q tables (chatcarlo/materials.py:376-384, 484-496)
Q_MAX = 20.0
N_Q = 2000
FLUOR_CUTOFF_KEV = 5.0  # chatcarlo/physics.py:38
N_LINES = 8

PART_PHOTO, PART_COMPT, PART_RAYL = 0, 1, 2


def energy_nodes() -> np.ndarray:
    """Node energies, uniform in ln E and rounded to float32-representable values."""
    e = np.exp(LN_E0 + DLN_E * np.arange(N_E)).astype(np.float32).astype(np.float64)
    e[0], e[-1] = E_MIN_KEV, E_MAX_KEV
    return e


def load_xaamdi(key: str):
    """E_keV, mu/rho, mu_en/rho of a bundled NIST XAAMDI table (chatcarlo/materials.py:298-306)."""
    data = np.loadtxt(DATA_DIR / "nist_xaamdi" / f"{key}.csv", delimiter=",", comments="#")
    return data[:, 0], data[:, 1], data[:, 2]


def load_hstar10():
    """ICRP-74 h*(10)/Phi table, E_keV and pSv cm^2 (chatcarlo/dose_coefficients.py:26-29)."""
    data = np.loadtxt(DATA_DIR / "h_star_10" / "photons_icrp74.csv", delimiter=",", comments="#")
    return data[:, 0], data[:, 1]


# ---------------------------------------------------------------------------
# per-element scattering cross sections from F and S
# ---------------------------------------------------------------------------
_QF = np.linspace(0.0, 14.0, 14001)   # fine q grid for integrals (q_max(150 keV)=12.1)


@functools.lru_cache(maxsize=None)
def _element_fs_fine(z: int):
    return A.form_factor(z, _QF), A.incoherent_sf(z, _QF)


@functools.lru_cache(maxsize=None)
def element_scatter_xs(z: int):
    """(sigma_coh, sigma_incoh) [cm^2/g] of element z on the energy nodes."""
    e = energy_nodes()
    f, s = _element_fs_fine(z)
    per_atom_to_mass = A.N_AVOGADRO / A.ATOMIC_WEIGHT[z]
    qmax = e / A.HC_KEV_ANGSTROM
    # coherent: pi r_e^2 (2/xm) int_0^xm (1 + mu^2) F^2 dx, mu = 1 - 2x/xm, x = q^2
    x = _QF ** 2
    f2 = f * f
    def cum(y):
        out = np.zeros_like(y)
        out[1:] = np.cumsum(0.5 * (y[1:] + y[:-1]) * np.diff(x))
        return out
    a0, a1, a2 = cum(f2), cum(f2 * x), cum(f2 * x * x)
    xm = qmax ** 2
    i0, i1, i2 = (np.interp(xm, x, a) for a in (a0, a1, a2))
    coh = math.pi * A.R_E_CM ** 2 * (2.0 / xm) * (2.0 * i0 - 4.0 * i1 / xm + 4.0 * i2 / xm ** 2)
    # incoherent: 2 pi int dsigma_KN/dOmega S(q) dmu, integrated over q (dmu = 4q dq / xm)
    incoh = np.empty_like(e)
    nq = 4001
    for j, ej in enumerate(e):
        q = np.linspace(0.0, qmax[j], nq)
        mu = 1.0 - 2.0 * q * q / xm[j]
        alpha = ej / 510.99895
        p = 1.0 / (1.0 + alpha * (1.0 - mu))
        kn = 0.5 * A.R_E_CM ** 2 * p * p * (p + 1.0 / p - (1.0 - mu * mu))
        sq = np.interp(q, _QF, s)
        y = 2.0 * math.pi * kn * sq * 4.0 * q / xm[j]
        incoh[j] = np.trapezoid(y, q)
    return coh * per_atom_to_mass, incoh * per_atom_to_mass


@functools.lru_cache(maxsize=None)
def element_compton_transfer(z: int):
    """Compton energy-transfer coefficient sigma_incoh <T>/E [cm^2/g] of element z on the nodes:
    the Klein-Nishina cross section weighted by S(Z,q) and by the share 1 - E'/E of the photon
    energy the electron takes, integrated over angle (the incoherent term of mu_tr)."""
    e = energy_nodes()
    _, s = _element_fs_fine(z)
    per_atom_to_mass = A.N_AVOGADRO / A.ATOMIC_WEIGHT[z]
    out = np.empty_like(e)
    nq = 4001
    for j, ej in enumerate(e):
        qmax = ej / A.HC_KEV_ANGSTROM
        xm = qmax * qmax
        q = np.linspace(0.0, qmax, nq)
        mu = 1.0 - 2.0 * q * q / xm
        alpha = ej / 510.99895
        p = 1.0 / (1.0 + alpha * (1.0 - mu))
        kn = 0.5 * A.R_E_CM ** 2 * p * p * (p + 1.0 / p - (1.0 - mu * mu))
        y = 2.0 * math.pi * kn * np.interp(q, _QF, s) * (1.0 - p) * 4.0 * q / xm
        out[j] = np.trapezoid(y, q)
    return out * per_atom_to_mass


def _segments(e_tab: np.ndarray):
    """Split a NIST energy column at its edges (consecutive energies within 1e-4 rel)."""
    cuts = [0]
    for i in range(1, len(e_tab)):
        if e_tab[i] / e_tab[i - 1] - 1.0 < 1e-4:
            cuts.append(i)
    cuts.append(len(e_tab))
    return [(cuts[k], cuts[k + 1]) for k in range(len(cuts) - 1)]


def _photo_from_nist(e_tab, mu_tab, coh_fn, incoh_fn):
    """Photo residual at NIST nodes, made positive by power-law continuation.

    Where NIST total - coherent - incoherent falls below 2 % of the total (light
    elements at high energy, where the residual is dominated by scatter-model
    error) the photo cross section is continued from the last reliable node
    with the local log-log slope clamped to [-3.5, -2.5].
    """
    res = mu_tab - coh_fn(e_tab) - incoh_fn(e_tab)
    out = res.copy()
    for lo, hi in _segments(e_tab):
        last = None
        slope = -3.0
        for i in range(lo, hi):
            if res[i] > 0.02 * mu_tab[i]:
                if last is not None and i - 1 >= lo and out[i - 1] > 0 and i - 1 == last:
                    s = math.log(out[i] / out[i - 1]) / math.log(e_tab[i] / e_tab[i - 1])
                    slope = min(-2.5, max(-3.5, s))
                last = i
            elif last is not None:
                out[i] = out[last] * (e_tab[i] / e_tab[last]) ** slope
            else:
                out[i] = max(res[i], 1e-3 * mu_tab[i])
    return np.maximum(out, 1e-12)


def _interp_segments(e_tab, val_tab, e_query):
    """Log-log linear interpolation of val_tab within edge-free segments.

    Returns values at e_query and the list of edges [(E_edge, v_below, v_above)].
    """
    segs = _segments(e_tab)
    le, lv = np.log(e_tab), np.log(val_tab)
    lq = np.log(e_query)
    out = np.empty_like(e_query)
    edges = []
    for si, (lo, hi) in enumerate(segs):
        e_lo = e_tab[lo]
        e_hi = e_tab[hi - 1]
        m = (e_query >= (e_lo if si > 0 else 0.0)) & (e_query < (e_tab[hi] if hi < len(e_tab) else np.inf))
        if not np.any(m):
            pass
        else:
            out[m] = np.exp(np.interp(lq[m], le[lo:hi], lv[lo:hi]))
        if si > 0:
            # edge between previous segment end and this segment start
            edges.append((float(e_tab[lo - 1]), float(val_tab[lo - 1]), float(val_tab[lo])))
    return out, edges


@functools.lru_cache(maxsize=None)
def element_photo(z: int):
    """Photo cross section [cm^2/g] of element z on the nodes, and its edges."""
    e = energy_nodes()
    if z in A.ELEMENT_TABLE:
        e_tab, mu_tab, _ = load_xaamdi(A.ELEMENT_TABLE[z])
        coh, incoh = element_scatter_xs(z)
        lne = np.log(e)
        coh_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(coh)))
        incoh_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(incoh)))
        photo_tab = _photo_from_nist(e_tab, mu_tab, coh_fn, incoh_fn)
        vals, edges = _interp_segments(e_tab, photo_tab, e)
        edges = [ed for ed in edges if E_MIN_KEV < ed[0] < E_MAX_KEV]
        return vals, edges
    # proxy: sigma_z(E) = sigma_ref(E * Ek_ref / Ek_z) * (z/zref)^4 * A_ref / A_z
    ref = A.PHOTO_PROXY[z]
    ref_vals_fn = _element_photo_fn(ref)
    scale_e = A.K_EDGE[ref] / A.K_EDGE[z]
    fac = (z / ref) ** 4 * A.ATOMIC_WEIGHT[ref] / A.ATOMIC_WEIGHT[z]
    vals = ref_vals_fn(e * scale_e) * fac
    edges = []
    for e_edge, vb, va in element_photo(ref)[1]:
        ez = e_edge / scale_e
        if E_MIN_KEV < ez < E_MAX_KEV:
            edges.append((ez, vb * fac, va * fac))
    return vals, edges


def _element_photo_fn(z: int):
    e_tab, mu_tab, _ = load_xaamdi(A.ELEMENT_TABLE[z])
    coh, incoh = element_scatter_xs(z)
    lne = np.log(energy_nodes())
    coh_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(coh)))
    incoh_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(incoh)))
    photo_tab = _photo_from_nist(e_tab, mu_tab, coh_fn, incoh_fn)

    def fn(ee):
        ee = np.clip(ee, e_tab[0], e_tab[-1])
        return _interp_segments(e_tab, photo_tab, np.atleast_1d(ee))[0]
    return fn


N_KFRAC = 12   # K-shell share nodes per element, uniform in ln E from the K edge to E_MAX_KEV


def _k_edge_segments(z: int):
    """Photo cross section of element z at the NIST nodes of the segments just below and just
    above its K edge: (E_below, photo_below, E_above, photo_above)."""
    e_tab, mu_tab, _ = load_xaamdi(A.ELEMENT_TABLE[z])
    coh, incoh = element_scatter_xs(z)
    lne = np.log(energy_nodes())
    coh_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(coh)))
    incoh_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(incoh)))
    photo_tab = _photo_from_nist(e_tab, mu_tab, coh_fn, incoh_fn)
    segs = _segments(e_tab)
    ek = A.K_SHELL[z]["edge"]
    for si in range(1, len(segs)):
        lo, hi = segs[si]
        if abs(e_tab[lo - 1] / ek - 1.0) < 0.01:
            blo, bhi = segs[si - 1]
            return e_tab[blo:bhi], photo_tab[blo:bhi], e_tab[lo:hi], photo_tab[lo:hi]
    raise ValueError(f"no K edge found in the NIST table of Z={z}")


def k_fraction_nodes(z: int):
    """K-shell share of photoabsorption above the K edge on N_KFRAC nodes E_j = E_K (E_MAX/E_K)^(j/11).

    The reference evaluates CS_Photo_Partial(Z,K,E)/CS_Photo(Z,E) per photon
    (chatcarlo/physics.py:193-197).  Here the share comes from the bundled NIST element table by
    edge-jump continuation: the photoabsorption of the outer shells (L, M, ...) is the cross
    section below the K edge continued above it along the quadratic in (ln E, ln sigma) through
    the last three NIST nodes below the edge (the outer-shell cross section keeps steepening
    as E moves away from its own edges), and the K share is 1 - sigma_nonK(E)/sigma(E), sigma(E)
    the photo cross section above the edge (log-log between the NIST nodes).  At the edge this
    is (J-1)/J, J the K-edge jump ratio.  Elements without a bundled table use their
    data-booklet jump ratio, constant in energy.
    Returns (nodes (N_KFRAC,), 1 / node spacing in ln E).
    """
    ks = A.K_SHELL[z]
    ek = float(ks["edge"])
    dln = math.log(E_MAX_KEV / ek) / (N_KFRAC - 1)
    e = ek * np.exp(dln * np.arange(N_KFRAC))
    if ks.get("jump") or z not in A.ELEMENT_TABLE:
        j = ks.get("jump") or (125.0 / z + 3.5)
        return np.full(N_KFRAC, (j - 1.0) / j), 1.0 / dln
    eb, pb, ea, pa = _k_edge_segments(z)
    c = np.polyfit(np.log(eb[-3:]), np.log(pb[-3:]), 2)
    non_k = pb[-1] * np.exp(np.polyval(c, np.log(e)) - np.polyval(c, math.log(eb[-1])))
    above = np.exp(np.interp(np.log(np.maximum(e, ea[0])), np.log(ea), np.log(pa)))
    return 1.0 - non_k / above, 1.0 / dln


def k_fraction(z: int, e_keV) -> np.ndarray:
    """K-shell share at energies e_keV >= the K edge: the nodes of k_fraction_nodes, linear in ln E."""
    k, inv = k_fraction_nodes(z)
    t = (np.log(np.asarray(e_keV, dtype=float)) - math.log(A.K_SHELL[z]["edge"])) * inv
    j = np.clip(np.floor(t).astype(int), 0, N_KFRAC - 2)
    return k[j] + (t - j) * (k[j + 1] - k[j])


# ---------------------------------------------------------------------------
# the pack
# ---------------------------------------------------------------------------
@dataclass
class DataPack:
    """All tables of the transport path.

    Material lookup tables hold, per material m, part p (photo, compton,
    rayleigh, mu_en, total = photo + compton + rayleigh) and node i: sig[m,p,i] and slope[m,p,i] =
    ln(sig[m,p,i+1]/sig[m,p,i]); the value between nodes is
    sig_i * exp(f * slope_i) with f = ln(E/E_i)/ln(E_{i+1}/E_i), i.e. log-log
    linear interpolation (chatcarlo/materials.py:240-253) written so that a
    float32 evaluation keeps ~1e-7 relative error.  Intervals holding an
    absorption edge carry an edge record for the photo, mu_en and total parts.
    """
    materials: list                     # names, index = material id
    density: np.ndarray                 # (n_mat,)
    e_node: np.ndarray                  # (N_E,) keV, float32-exact
    dl_node: np.ndarray                 # (N_E-1,) ln(E_{i+1}/E_i)
    sig: np.ndarray                     # (n_mat, 5, N_E) cm^2/g: photo, compt, rayl, muen, total
    slope: np.ndarray                   # (n_mat, 5, N_E-1)
    edge_of_interval: np.ndarray        # (n_mat, N_E) int32, -1 = none
    edge_e: np.ndarray                  # (n_mat, MAX_EDGES) keV
    edge_dl: np.ndarray                 # (n_mat, MAX_EDGES, 2) ln(Ee/Ei), ln(Ei+1/Ee)
    edge_sig: np.ndarray                # (n_mat, MAX_EDGES, 3 part, 2 side): photo, muen, total
    edge_slope: np.ndarray              # (n_mat, MAX_EDGES, 3 part, 2 side)
    mat_elem: list                      # per material: list of element ids
    mat_frac: list                      # per material: mass fractions
    sel_cdf: list                       # per material: (3, N_E, n_elem) cumulative fractions
    elements: list                      # Z per element id
    sq_over_z: np.ndarray               # (n_el, N_Q) S/Z on q_s grid
    q_s: np.ndarray                     # (N_Q,)
    ray_x: np.ndarray                   # (N_Q,) x = q^2 grid
    ray_a: np.ndarray                   # (n_el, N_Q) cumulative A(x)
    fl_edge: np.ndarray                 # (n_el,) K edge keV (0 = no fluorescence)
    fl_omega: np.ndarray                # (n_el,)
    fl_kfrac: np.ndarray                # (n_el, N_KFRAC) K-shell share nodes, uniform in ln E from the edge
    fl_kinv: np.ndarray                 # (n_el,) 1 / node spacing of fl_kfrac in ln E
    fl_lines: np.ndarray                # (n_el, 8) keV
    fl_cdf: np.ndarray                  # (n_el, 8) cumulative line probabilities
    h10_e: np.ndarray                   # (25,) keV
    h10_v: np.ndarray                   # (25,) pSv cm^2
    h10_dl: np.ndarray                  # (24,) ln(E_{j+1}/E_j)
    h10_slope: np.ndarray               # (24,) ln(h_{j+1}/h_j)
    dtype: str = "f64"

    def mat_id(self, name: str) -> int:
        key = name.lower().strip()
        key = A.SYMBOL_MATERIAL.get(key, key)
        if key not in self.materials:
            raise ValueError(f"unknown material '{name}'. candidates: {sorted(self.materials)}")
        return self.materials.index(key)

    def rounded(self) -> "DataPack":
        """The pack as the device sees it: every float table rounded to float32."""
        kw = {}
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray) and v.dtype.kind == "f":
                kw[k] = v.astype(np.float32).astype(np.float64)
            elif k in ("mat_frac", "sel_cdf"):
                kw[k] = [np.asarray(x, dtype=np.float32).astype(np.float64) for x in v]
            elif isinstance(v, np.ndarray):
                kw[k] = v.copy()
            elif isinstance(v, list):
                kw[k] = [list(x) if isinstance(x, list) else x for x in v]
            else:
                kw[k] = v
        kw["dtype"] = "f32"
        return DataPack(**kw)


# Relative errors assigned to the two photoabsorption estimates combined in _material_photo:
# the atomic model's scatter cross sections (water at 60 keV: incoherent +0.4 %, coherent +0.7 %
# against the reference's EPDL values), the Compton transfer coefficient, NIST mu_en/rho, and the
# K-fluorescence escape correction.
ERR_SCATTER = 5e-3
ERR_TRANSFER = 3e-3
ERR_MUEN = 2e-3
ERR_ESCAPE = 0.2


def _k_escape(comp, e_tab, photo_el):
    """Share of absorbed photon energy carried off by K fluorescence, per NIST node: photo-weighted
    over the elements, k_share x omega_K x mean K-line energy / E above each K edge (the term by
    which mu_en/rho falls short of photo + transfer).  photo_el: per element photo at e_tab."""
    tot = sum(w * p for (z, w), p in zip(comp, photo_el))
    x = np.zeros_like(e_tab)
    for (z, w), p in zip(comp, photo_el):
        ek = A.K_EDGE.get(z, 0.0)
        if ek <= 0.0:
            continue
        above = e_tab > ek * (1.0 + 1e-6)
        if not np.any(above):
            continue
        ks = A.K_SHELL.get(z)
        if ks is not None:
            rates = np.asarray(ks["rates"], dtype=float)
            e_line = float(np.dot(rates, ks["lines"]) / rates.sum())
            share = k_fraction(z, e_tab[above]) if z in A.ELEMENT_TABLE and not ks.get("jump") else \
                (ks["jump"] - 1.0) / ks["jump"]
        else:
            e_line = 0.9 * ek
            j = 125.0 / z + 3.5
            share = (j - 1.0) / j
        x[above] += w * p[above] / tot[above] * share * A.k_fluorescence_yield(z) * e_line / e_tab[above]
    return x


def _material_photo(comp, e_tab, mu_tab, muen_tab, coh_fn, incoh_fn, tr_fn):
    """Photoabsorption [cm^2/g] at the NIST nodes of a material, from two estimates:

    * residual: NIST total - coherent - incoherent of the atomic model; its error is the scatter
      model's, ERR_SCATTER x (coherent + incoherent), large where scatter dominates;
    * energy absorption: mu_en/rho = photo (1 - X) + transfer, the relation by which NIST's
      mu_en/rho is built (Hubbell & Seltzer), so photo = (mu_en/rho - transfer) / (1 - X), X the
      K-fluorescence escape share (_k_escape); L-shell fluorescence is not modelled, so this
      estimate is used only where elements above Z = 26 take under 1 % of the photoabsorption.

    The two are combined with inverse-variance weights per node.  For water at 60 keV the result
    is 0.01491 cm^2/g against the reference's EPDL 0.014926 (the residual alone gives 0.01412).
    """
    res = _photo_from_nist(e_tab, mu_tab, coh_fn, incoh_fn)
    photo_el = [_element_photo_fn(z)(e_tab) if z in A.ELEMENT_TABLE else element_photo_at(z, e_tab)
                for z, _ in comp]
    tot_el = sum(w * p for (z, w), p in zip(comp, photo_el))
    heavy = sum(w * p for (z, w), p in zip(comp, photo_el) if z > 26) / tot_el
    x = _k_escape(comp, e_tab, photo_el)
    tr = tr_fn(e_tab)
    men = (muen_tab - tr) / (1.0 - x)
    var_res = (ERR_SCATTER * (coh_fn(e_tab) + incoh_fn(e_tab))) ** 2
    var_men = (ERR_TRANSFER * tr) ** 2 + (ERR_MUEN * muen_tab) ** 2 + (ERR_ESCAPE * x * men) ** 2
    use = (heavy < 0.01) & (men > 0.0)
    w_men = np.where(use, var_res / (var_res + var_men), 0.0)
    return np.maximum((1.0 - w_men) * res + w_men * np.where(use, men, 0.0), 1e-12)


def element_photo_at(z: int, e_q):
    """Photo cross section of element z at energies e_q (log-log between the pack's nodes)."""
    vals, _ = element_photo(z)
    return np.exp(np.interp(np.log(e_q), np.log(energy_nodes()), np.log(vals)))


def _material_tables(name: str):
    """Node values and edge records of one material (photo, compton, rayleigh, mu_en).

    Photoabsorption from _material_photo; coherent and incoherent from the atomic model, scaled
    together per node so that the three parts add up to NIST's total."""
    dens, comp, key = A.MATERIALS[name]
    e = energy_nodes()
    coh = np.zeros_like(e)
    incoh = np.zeros_like(e)
    tr = np.zeros_like(e)
    for z, w in comp:
        c, i = element_scatter_xs(z)
        coh += w * c
        incoh += w * i
        tr += w * element_compton_transfer(z)
    e_tab, mu_tab, muen_tab = load_xaamdi(key)
    lne = np.log(e)
    coh_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(coh)))
    incoh_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(incoh)))
    tr_fn = lambda ee: np.exp(np.interp(np.log(ee), lne, np.log(tr)))
    photo_tab = _material_photo(comp, e_tab, mu_tab, muen_tab, coh_fn, incoh_fn, tr_fn)
    photo, _ = _interp_segments(e_tab, photo_tab, e)
    # scatter closes the NIST total: mu_rho(E) at the nodes, log-log between NIST nodes
    mu_nodes, _ = _interp_segments(e_tab, mu_tab, e)
    scale = (mu_nodes - photo) / (coh + incoh)
    scale = np.where(scale > 0.0, scale, 1.0)
    coh = coh * scale
    incoh = incoh * scale
    # mu_en: PCHIP in log-log space through every NIST node (chatcarlo/materials.py:309-312)
    pch = PchipInterpolator(np.log(e_tab), np.log(muen_tab))
    muen = np.exp(pch(lne))
    edge_rec = []
    segs = _segments(e_tab)
    for si in range(1, len(segs)):
        lo = segs[si][0]
        ee = float(np.float32(e_tab[lo - 1]))
        if not (E_MIN_KEV < ee < E_MAX_KEV):
            continue
        edge_rec.append((ee, photo_tab[lo - 1], photo_tab[lo],
                         float(muen_tab[lo - 1]), float(muen_tab[lo])))
    return dens, comp, np.stack([photo, incoh, coh, muen, photo + incoh + coh]), edge_rec


@functools.lru_cache(maxsize=1)
def build_pack() -> DataPack:
    names = list(A.MATERIALS)
    n_mat = len(names)
    elements = sorted({z for nm in names for z, _ in A.MATERIALS[nm][1]})
    el_index = {z: i for i, z in enumerate(elements)}
    e = energy_nodes()
    dl_node = np.log(e[1:] / e[:-1])
    sig = np.zeros((n_mat, 5, N_E))
    edge_of_interval = np.full((n_mat, N_E), -1, dtype=np.int32)
    edge_e = np.zeros((n_mat, MAX_EDGES))
    edge_dl = np.zeros((n_mat, MAX_EDGES, 2))
    edge_sig = np.zeros((n_mat, MAX_EDGES, 3, 2))
    edge_slope = np.zeros((n_mat, MAX_EDGES, 3, 2))
    dens = np.zeros(n_mat)
    mat_elem, mat_frac, sel_cdf = [], [], []
    for m, nm in enumerate(names):
        d, comp, parts, edges = _material_tables(nm)
        dens[m] = d
        sig[m] = parts
        assert len(edges) <= MAX_EDGES
        for k, (ee, pb, pa, mb, ma) in enumerate(edges):
            iv = int(np.searchsorted(e, ee, side="right")) - 1
            iv = min(max(iv, 0), N_E - 2)
            assert edge_of_interval[m, iv] < 0, f"two edges in one interval for {nm}"
            assert e[iv] < ee < e[iv + 1]
            edge_of_interval[m, iv] = k
            edge_e[m, k] = ee
            edge_dl[m, k] = (math.log(ee / e[iv]), math.log(e[iv + 1] / ee))
            # smooth parts at the edge energy, log-log between the two nodes
            fe = math.log(ee / e[iv]) / math.log(e[iv + 1] / e[iv])
            smooth = sum(parts[p, iv] * (parts[p, iv + 1] / parts[p, iv]) ** fe for p in (1, 2))
            recs = ((pb, pa, 0), (mb, ma, 3), (pb + smooth, pa + smooth, 4))
            for p, (vb, va, part) in enumerate(recs):
                edge_sig[m, k, p] = (vb, va)
                edge_slope[m, k, p] = (math.log(vb / parts[part, iv]),
                                       math.log(parts[part, iv + 1] / va))
        # element selection: w_i sigma_type,i / sum (chatcarlo/materials.py:358-440)
        zs = [z for z, _ in comp]
        ws = np.array([w for _, w in comp])
        per = np.zeros((3, len(zs), N_E))
        for k, z in enumerate(zs):
            c, i = element_scatter_xs(z)
            p, _ = element_photo(z)
            per[0, k], per[1, k], per[2, k] = p, i, c
        weighted = ws[None, :, None] * per
        tot = weighted.sum(axis=1, keepdims=True)
        tot = np.where(tot > 0, tot, 1.0)
        cdf = np.cumsum(weighted / tot, axis=1)           # (3, n_elem, N_E)
        cdf[:, -1, :] = 1.0
        sel_cdf.append(np.ascontiguousarray(np.transpose(cdf, (0, 2, 1))))  # (3, N_E, n_elem)
        mat_elem.append([el_index[z] for z in zs])
        mat_frac.append(ws)
    slope = np.log(sig[:, :, 1:] / sig[:, :, :-1])
    # per-element q tables (chatcarlo/materials.py:376-406, 484-496)
    n_el = len(elements)
    q_s = np.concatenate([[0.0], np.linspace(1e-3, Q_MAX, N_Q - 1)])
    q_r = np.linspace(0.0, Q_MAX, N_Q)
    ray_x = q_r ** 2
    sq = np.zeros((n_el, N_Q))
    ray_a = np.zeros((n_el, N_Q))
    fl_edge = np.zeros(n_el)
    fl_omega = np.zeros(n_el)
    fl_kfrac = np.zeros((n_el, N_KFRAC))
    fl_kinv = np.zeros(n_el)
    fl_lines = np.zeros((n_el, N_LINES))
    fl_cdf = np.zeros((n_el, N_LINES))
    for i, z in enumerate(elements):
        s = A.incoherent_sf(z, q_s)
        s[0] = 0.0
        sq[i] = s / z
        f = A.form_factor(z, q_r)
        y = f * f
        ray_a[i, 1:] = np.cumsum(0.5 * (y[1:] + y[:-1]) * np.diff(ray_x))
        ks = A.K_SHELL.get(z)
        if ks is not None:
            lines = np.array(ks["lines"], dtype=float)
            rates = np.array(ks["rates"], dtype=float)
            rates = rates / rates.sum()
            if lines.max() >= FLUOR_CUTOFF_KEV:
                fl_edge[i] = ks["edge"]
                fl_omega[i] = ks["omega"]
                fl_kfrac[i], fl_kinv[i] = k_fraction_nodes(z)
                fl_lines[i] = lines
                fl_cdf[i] = np.cumsum(rates)
                fl_cdf[i, -1] = 1.0
    h10_e, h10_v = load_hstar10()
    return DataPack(materials=names, density=dens, e_node=e, dl_node=dl_node, sig=sig,
                    slope=slope, edge_of_interval=edge_of_interval, edge_e=edge_e,
                    edge_dl=edge_dl, edge_sig=edge_sig, edge_slope=edge_slope,
                    mat_elem=mat_elem, mat_frac=mat_frac, sel_cdf=sel_cdf, elements=elements,
                    sq_over_z=sq, q_s=q_s, ray_x=ray_x, ray_a=ray_a, fl_edge=fl_edge,
                    fl_omega=fl_omega, fl_kfrac=fl_kfrac, fl_kinv=fl_kinv, fl_lines=fl_lines, fl_cdf=fl_cdf,
                    h10_e=h10_e, h10_v=h10_v, h10_dl=np.log(h10_e[1:] / h10_e[:-1]),
                    h10_slope=np.log(h10_v[1:] / h10_v[:-1]))
