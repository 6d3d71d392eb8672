"""Host-side material queries over the data pack (set-up and reporting only).

Mirrors the reference's public lookups (chatcarlo/materials.py:292-340,
499-518): mu_rho, mu_rho_parts, mu_en_rho, linear_mu, density.  They serve
host set-up (the Kramers spectrum's aluminium filter, dose conversion,
reporting); photon transport evaluates the same tables on the GPU.
"""
from __future__ import annotations

import numpy as np

from .datapack import build_pack

MATERIAL_ALIASES = {"aluminium": "aluminum"}


def _pack():
    return build_pack()


def resolve(material: str) -> int:
    key = material.lower().strip()
    key = MATERIAL_ALIASES.get(key, key)
    return _pack().mat_id(key)


def density(material: str) -> float:
    return float(_pack().density[resolve(material)])


def _lookup(material: str, energies_keV, part: int):
    p = _pack()
    m = resolve(material)
    e = np.atleast_1d(np.asarray(energies_keV, dtype=float))
    if e.min() < p.e_node[0] or e.max() > p.e_node[-1]:
        raise ValueError(f"energy {e.min():.3g}-{e.max():.3g} keV outside the cross-section "
                         f"table range [{p.e_node[0]:.3g}, {p.e_node[-1]:.3g}] keV")
    i = np.clip(np.searchsorted(p.e_node, e, side="right") - 1, 0, len(p.e_node) - 2)
    f = np.log(e / p.e_node[i]) / p.dl_node[i]
    val = p.sig[m, part, i] * np.exp(f * p.slope[m, part, i])
    ep = {0: 0, 3: 1, 4: 2}.get(part)
    if ep is not None:
        k = p.edge_of_interval[m, i]
        for j in np.where(k >= 0)[0]:
            kk, ii = k[j], i[j]
            if e[j] < p.edge_e[m, kk]:
                fb = np.log(e[j] / p.e_node[ii]) / p.edge_dl[m, kk, 0]
                val[j] = p.sig[m, part, ii] * np.exp(fb * p.edge_slope[m, kk, ep, 0])
            else:
                fa = np.log(e[j] / p.edge_e[m, kk]) / p.edge_dl[m, kk, 1]
                val[j] = p.edge_sig[m, kk, ep, 1] * np.exp(fa * p.edge_slope[m, kk, ep, 1])
    return val


def mu_rho_parts(material: str, energies_keV) -> dict:
    """photoelectric / compton / rayleigh mass attenuation [cm^2/g]."""
    return {"photoelectric": _lookup(material, energies_keV, 0),
            "compton": _lookup(material, energies_keV, 1),
            "rayleigh": _lookup(material, energies_keV, 2)}


def mu_rho(material: str, energies_keV) -> np.ndarray:
    parts = mu_rho_parts(material, energies_keV)
    return parts["photoelectric"] + parts["compton"] + parts["rayleigh"]


def mu_en_rho(material: str, energies_keV) -> np.ndarray:
    return _lookup(material, energies_keV, 3)


def linear_mu(material: str, energies_keV) -> np.ndarray:
    return mu_rho(material, energies_keV) * density(material)
