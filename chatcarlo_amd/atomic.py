"""Atomic data model for the photon cross-section data pack.

The reference builds every per-element table from xraylib (EPDL97) at run time
(chatcarlo/materials.py:202-225 cross sections, :375-406 F(Z,q), :483-496
S(Z,q), :443-480 K-shell data, :343-355 compositions). xraylib is not
installable offline, so this module supplies a self-contained replacement:

* Atomic form factor F(Z,q) and incoherent scattering function S(Z,q) from
  Slater-type orbitals whose effective charges follow Slater's screening
  rules (J. C. Slater, Phys. Rev. 36, 57 (1930)).  For one electron in a
  shell with effective quantum number n* and exponent zeta = Z_eff / n*, the
  density is rho(r) ~ r^(2n*-2) exp(-2 zeta r) and its Fourier transform is

      f(k) = beta * (beta^2 / (beta^2 + k^2))^n* * sin(2 n* atan(k/beta)) / (2 n* k),

  with beta = 2 zeta (bohr^-1) and k = 4 pi a0 q (q = sin(theta/2)/lambda in
  1/Angstrom, the xraylib MomentTransf convention used by the reference,
  chatcarlo/physics.py:35,122).  F = sum_shell N_e f and, in the
  independent-electron Waller-Hartree approximation without exchange,
  S = Z - sum_shell N_e f^2.  For hydrogen this reduces to the exact
  F = (1 + k^2/4)^-2.
* K-shell fluorescence data (edge, omega_K, eight K lines and their relative
  radiative rates) for the elements whose K lines reach the 5 keV cut-off of
  chatcarlo/physics.py:38; line energies and rates are standard X-ray data-booklet values, the
  fluorescence yields those of the Bambynek fit (k_fluorescence_yield).
* Material compositions (mass fractions) and densities of the NIST compound
  data base entries the reference aliases (chatcarlo/materials.py:34-49).
"""
from __future__ import annotations

import math

import numpy as np

A0_ANGSTROM = 0.529177210903
R_E_CM = 2.8179403262e-13
N_AVOGADRO = 6.02214076e23
HC_KEV_ANGSTROM = 12.3984193   # chatcarlo/physics.py:35
MEC2_KEV = 511.0               # chatcarlo/physics.py:34

SYMBOL = {1: "H", 6: "C", 7: "N", 8: "O", 11: "Na", 12: "Mg", 13: "Al", 14: "Si",
          15: "P", 16: "S", 17: "Cl", 18: "Ar", 19: "K", 20: "Ca", 22: "Ti", 26: "Fe",
          29: "Cu", 30: "Zn", 33: "As", 74: "W", 82: "Pb"}
Z_OF = {v: k for k, v in SYMBOL.items()}

ATOMIC_WEIGHT = {1: 1.008, 6: 12.011, 7: 14.007, 8: 15.999, 11: 22.98977, 12: 24.305,
                 13: 26.98154, 14: 28.0855, 15: 30.97376, 16: 32.06, 17: 35.453,
                 18: 39.948, 19: 39.0983, 20: 40.078, 22: 47.867, 26: 55.845, 29: 63.546,
                 30: 65.38, 33: 74.9216, 74: 183.84, 82: 207.2}

# Ground-state occupations grouped as Slater groups: (n, kind, electrons),
# kind "sp" for (ns,np), "d", "f".
_CORE_NE = [(1, "sp", 2), (2, "sp", 8)]
_CORE_AR = _CORE_NE + [(3, "sp", 8)]
_CORE_KR = _CORE_AR[:-1] + [(3, "sp", 8), (3, "d", 10), (4, "sp", 8)]
_CORE_XE = _CORE_KR + [(4, "d", 10), (5, "sp", 8)]
CONFIG = {
    1: [(1, "sp", 1)],
    6: [(1, "sp", 2), (2, "sp", 4)],
    7: [(1, "sp", 2), (2, "sp", 5)],
    8: [(1, "sp", 2), (2, "sp", 6)],
    11: _CORE_NE + [(3, "sp", 1)],
    12: _CORE_NE + [(3, "sp", 2)],
    13: _CORE_NE + [(3, "sp", 3)],
    14: _CORE_NE + [(3, "sp", 4)],
    15: _CORE_NE + [(3, "sp", 5)],
    16: _CORE_NE + [(3, "sp", 6)],
    17: _CORE_NE + [(3, "sp", 7)],
    18: _CORE_NE + [(3, "sp", 8)],
    19: _CORE_AR + [(4, "sp", 1)],
    20: _CORE_AR + [(4, "sp", 2)],
    22: _CORE_AR + [(3, "d", 2), (4, "sp", 2)],
    26: _CORE_AR + [(3, "d", 6), (4, "sp", 2)],
    29: _CORE_AR + [(3, "d", 10), (4, "sp", 1)],
    30: _CORE_AR + [(3, "d", 10), (4, "sp", 2)],
    33: _CORE_AR + [(3, "d", 10), (4, "sp", 5)],
    74: _CORE_XE + [(4, "f", 14), (5, "d", 4), (6, "sp", 2)],
    82: _CORE_XE + [(4, "f", 14), (5, "d", 10), (6, "sp", 4)],
}
_NSTAR = {1: 1.0, 2: 2.0, 3: 3.0, 4: 3.7, 5: 4.0, 6: 4.2}
# Slater ordering of groups: (1s)(2sp)(3sp)(3d)(4sp)(4d)(4f)(5sp)(5d)(5f)(6sp)
_ORDER = [(1, "sp"), (2, "sp"), (3, "sp"), (3, "d"), (4, "sp"), (4, "d"), (4, "f"),
          (5, "sp"), (5, "d"), (5, "f"), (6, "sp")]


def _sorted_groups(z: int):
    cfg = {(n, k): e for n, k, e in CONFIG[z]}
    assert sum(cfg.values()) == z, (z, sum(cfg.values()))
    return [(n, k, cfg[(n, k)]) for n, k in _ORDER if (n, k) in cfg]


def slater_shells(z: int):
    """[(n*, Z_eff, electrons)] per Slater group of element z (Slater 1930 rules)."""
    groups = _sorted_groups(z)
    out = []
    for gi, (n, kind, ne) in enumerate(groups):
        s = 0.0
        same = 0.30 if n == 1 else 0.35
        s += same * (ne - 1)
        for gj, (n2, kind2, ne2) in enumerate(groups):
            if gj >= gi:
                continue
            if kind == "sp":
                if n2 == n - 1:
                    s += 0.85 * ne2
                elif n2 <= n - 2:
                    s += 1.00 * ne2
                else:
                    # same n, lower group cannot precede an (ns,np) group
                    s += 1.00 * ne2
            else:
                s += 1.00 * ne2
        out.append((_NSTAR[n], z - s, ne))
    return out


def _orbital_ff(k_au: np.ndarray, nstar: float, zeff: float) -> np.ndarray:
    beta = 2.0 * zeff / nstar
    k = np.asarray(k_au, dtype=float)
    out = np.ones_like(k)
    m = k > 0
    km = k[m]
    ratio = beta * beta / (beta * beta + km * km)
    phi = np.arctan2(km, beta)
    out[m] = beta * ratio ** nstar * np.sin(2.0 * nstar * phi) / (2.0 * nstar * km)
    return out


def q_to_k_au(q_inv_angstrom):
    return 4.0 * math.pi * A0_ANGSTROM * np.asarray(q_inv_angstrom, dtype=float)


def form_factor(z: int, q) -> np.ndarray:
    """F(Z,q), q in 1/Angstrom (xraylib MomentTransf convention)."""
    k = q_to_k_au(q)
    f = np.zeros_like(k)
    for nstar, zeff, ne in slater_shells(z):
        f += ne * _orbital_ff(k, nstar, zeff)
    return f


def incoherent_sf(z: int, q) -> np.ndarray:
    """S(Z,q) = Z - sum_i f_i(q)^2 (independent electrons, no exchange)."""
    k = q_to_k_au(q)
    s = np.full_like(k, float(z))
    for nstar, zeff, ne in slater_shells(z):
        s -= ne * _orbital_ff(k, nstar, zeff) ** 2
    return np.clip(s, 0.0, float(z))


# ---- K-shell fluorescence (elements with K lines >= 5 keV) -----------------
def k_fluorescence_yield(z: int) -> float:
    """omega_K from the fit of W. Bambynek (1984) recommended by Hubbell et al., J. Phys. Chem.
    Ref. Data 23, 339 (1994): s = sum_i C_i Z^i, omega_K = s^4 / (1 + s^4).  It reproduces the
    yields the reference logs from xraylib (Cu 0.454, W 0.954, Pb 0.9634)."""
    s = 3.70e-2 + 3.112e-2 * z + 5.44e-5 * z * z - 1.25e-6 * z ** 3
    return s ** 4 / (1.0 + s ** 4)



# line order: KL2, KL3, KM2, KM3, KN2, KN3, KO, KP (chatcarlo/materials.py:459-460)
# energies keV; rates are relative radiative rates (normalised on use).
K_SHELL = {
    26: dict(edge=7.112, omega=k_fluorescence_yield(26), jump=None,
             lines=[6.39084, 6.40384, 7.05798, 7.05798, 0, 0, 0, 0],
             rates=[0.2976, 0.5836, 0.0396, 0.0792, 0, 0, 0, 0]),
    29: dict(edge=8.979, omega=k_fluorescence_yield(29), jump=None,
             lines=[8.02783, 8.04778, 8.90529, 8.90529, 0, 0, 0, 0],
             rates=[0.2970, 0.5830, 0.0400, 0.0800, 0, 0, 0, 0]),
    30: dict(edge=9.659, omega=k_fluorescence_yield(30), jump=7.67,
             lines=[8.61578, 8.63886, 9.572, 9.572, 0, 0, 0, 0],
             rates=[0.2965, 0.5815, 0.0405, 0.0815, 0, 0, 0, 0]),
    33: dict(edge=11.867, omega=k_fluorescence_yield(33), jump=7.1,
             lines=[10.50799, 10.54372, 11.7203, 11.7262, 11.8643, 11.8643, 0, 0],
             rates=[0.2860, 0.5560, 0.0480, 0.0940, 0.0050, 0.0110, 0, 0]),
    74: dict(edge=69.525, omega=k_fluorescence_yield(74), jump=None,
             lines=[57.9817, 59.3182, 66.9514, 67.2443, 69.0300, 69.1010, 69.4500, 69.5200],
             rates=[0.2880, 0.5000, 0.0550, 0.1070, 0.0140, 0.0270, 0.0080, 0.0010]),
    82: dict(edge=88.004, omega=k_fluorescence_yield(82), jump=None,
             lines=[72.8042, 74.9694, 84.4500, 84.9360, 87.3000, 87.3670, 87.9100, 87.9900],
             rates=[0.2928, 0.4924, 0.0570, 0.1093, 0.0135, 0.0263, 0.0075, 0.0012]),
}

# Nominal K edges of the remaining elements (keV); used to scale photo cross
# sections of elements without a bundled NIST element table.
K_EDGE = {1: 0.0136, 6: 0.2838, 7: 0.4016, 8: 0.5320, 11: 1.0721, 12: 1.3050, 13: 1.5596,
          14: 1.8389, 15: 2.1455, 16: 2.4720, 17: 2.8224, 18: 3.2029, 19: 3.6074,
          20: 4.0381, 22: 4.9664, 26: 7.112, 29: 8.979, 30: 9.659, 33: 11.867, 74: 69.525, 82: 88.004}

# ---- materials --------------------------------------------------------------
# name -> (density g/cm^3, [(Z, mass fraction)], NIST XAAMDI table key)
MATERIALS = {
    "water": (1.0, [(1, 0.111894), (8, 0.888106)], "water"),
    "air": (1.20479e-3, [(6, 0.000124), (7, 0.755268), (8, 0.231781), (18, 0.012827)], "air"),
    "soft_tissue": (1.0, [(1, 0.104472), (6, 0.232190), (7, 0.024880), (8, 0.630238),
                          (11, 0.001130), (12, 0.000130), (15, 0.001330), (16, 0.001990),
                          (17, 0.001340), (19, 0.001990), (20, 0.000230), (26, 0.000050),
                          (30, 0.000030)], "soft_tissue"),
    "bone": (1.85, [(1, 0.047234), (6, 0.144330), (7, 0.041990), (8, 0.446096),
                    (12, 0.002200), (15, 0.104970), (16, 0.003150), (20, 0.209930),
                    (30, 0.000100)], "bone"),
    "lung": (1.05, [(1, 0.101278), (6, 0.102310), (7, 0.028650), (8, 0.757072),
                    (11, 0.001840), (12, 0.000730), (15, 0.000800), (16, 0.002250),
                    (17, 0.002660), (19, 0.001940), (20, 0.000090), (26, 0.000370),
                    (30, 0.000010)], "lung"),
    "pmma": (1.19, [(1, 0.080538), (6, 0.599848), (8, 0.319614)], "pmma"),
    "concrete": (2.3, [(1, 0.010000), (6, 0.001000), (8, 0.529107), (11, 0.016000),
                       (12, 0.002000), (13, 0.033872), (14, 0.337021), (19, 0.013000),
                       (20, 0.044000), (26, 0.014000)], "concrete"),
    "lead": (11.35, [(82, 1.0)], "z82_Pb"),
    "aluminum": (2.699, [(13, 1.0)], "z13_Al"),
    "copper": (8.96, [(29, 1.0)], "z29_Cu"),
    "iron": (7.874, [(26, 1.0)], "z26_Fe"),
    "tungsten": (19.3, [(74, 1.0)], "z74_W"),
    "calcium": (1.55, [(20, 1.0)], "z20_Ca"),
    # ICRU Report 44 compositions as tabulated by NIST XAAMDI (Hubbell & Seltzer, Table 2)
    "muscle": (1.05, [(1, 0.102), (6, 0.143), (7, 0.034), (8, 0.710), (11, 0.001), (15, 0.002),
                      (16, 0.003), (17, 0.001), (19, 0.004)], "muscle"),
    "adipose": (0.95, [(1, 0.114), (6, 0.598), (7, 0.007), (8, 0.278), (11, 0.001), (16, 0.001),
                       (17, 0.001)], "adipose"),
    # NIST compound "Glass, Lead" (the reference's lead_glass alias, chatcarlo/materials.py:46)
    "lead_glass": (6.22, [(8, 0.156453), (14, 0.080866), (22, 0.008092), (33, 0.002651),
                          (82, 0.751938)], "lead_glass"),
}
# Element symbols resolve to the element materials (chatcarlo/materials.py:52-65 resolves any
# symbol through xraylib; here the symbols of the bundled element tables).
SYMBOL_MATERIAL = {"pb": "lead", "al": "aluminum", "cu": "copper", "fe": "iron", "w": "tungsten",
                   "ca": "calcium"}
# Element tables bundled from NIST XAAMDI (chatcarlo/data/nist_xaamdi/z*.csv).
ELEMENT_TABLE = {1: "z01_H", 6: "z06_C", 7: "z07_N", 8: "z08_O", 13: "z13_Al",
                 20: "z20_Ca", 26: "z26_Fe", 29: "z29_Cu", 74: "z74_W", 82: "z82_Pb"}
# Element without its own table -> tabulated element whose photo cross section
# is scaled to it (edge-energy shifted, Z^4/A scaled).
PHOTO_PROXY = {11: 13, 12: 13, 14: 13, 15: 13, 16: 20, 17: 20, 18: 20, 19: 20, 22: 20, 30: 29, 33: 29}
