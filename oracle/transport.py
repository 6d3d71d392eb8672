"""Oracle batch transport: the reference's round-based analog loop in FP64 NumPy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Also the timed CPU
baseline of bench.py ("the reference NumPy path" restated on the same data pack).

Restates chatcarlo/transport.py:98-233 (transport_photons) and :257-312
(_run_batches): one shared PCG64 stream per batch, tau = -ln(U), material at
the current point, step to the next analytic boundary or the interaction
point, nudge 1e-6 cm past boundaries, interaction type by the part
fractions, photo + K fluorescence / bound Compton / Rayleigh, collision
estimator per material, track-length kerma and H*(10) per segment, batch
moments per batch.  Pinned by Beer-Lambert (reference tests/test_transport.py:30-51),
energy conservation (tests/test_fluorescence.py:80-97) and collision vs
track-length agreement (tests/test_tally.py:54-75), restated in
tests/test_cpu_host.py (test_oracle_beer_lambert, test_oracle_energy_conservation,
test_oracle_collision_vs_tracklength).
"""
from __future__ import annotations

import numpy as np

from . import core
from .source import sample_source


class Grid:
    """Tally arrays with the reference's batch moments (chatcarlo/tally.py:26-94)."""

    def __init__(self, origin, shape, voxel, track_uncertainty=True):
        self.origin = np.asarray(origin, dtype=float)
        self.shape = tuple(int(x) for x in shape)
        self.voxel = float(voxel)
        self.kerma = np.zeros(self.shape)
        self.h10 = np.zeros(self.shape)
        self.track = track_uncertainty
        self.kerma_q = np.zeros(self.shape)
        self.h10_q = np.zeros(self.shape)
        self.hits = np.zeros(self.shape, dtype=np.int32)
        self._pk = np.zeros(self.shape)
        self._ph = np.zeros(self.shape)
        self.n_batches = 0
        self.n_histories = 0

    def end_batch(self, nb):
        if not self.track or nb <= 0:
            return
        dk = self.kerma - self._pk
        dh = self.h10 - self._ph
        self.kerma_q += dk ** 2 / nb
        self.h10_q += dh ** 2 / nb
        self.hits += ((dk != 0) | (dh != 0)).astype(np.int32)
        np.copyto(self._pk, self.kerma)
        np.copyto(self._ph, self.h10)
        self.n_batches += 1
        self.n_histories += nb


def _draw_gen(rng):
    return lambda idx, slot: rng.random(len(idx))


EVENTS = ("boundary", "escape", "photoelectric", "fluorescence", "compton", "rayleigh")


def transport_photons(pos, dirv, energy, geom, pack, rng, grid=None, fluorescence=True,
                      exact_tally=True, tally_rng=None, k_trials=100000, region=None, recorder=None):
    """chatcarlo/transport.py:98-233 on the oracle's geometry and pack.

    region = (centre, radius): adds the per-history H*(10) track-length score
    in that sphere, h*(10)(E) x chord / volume, as result["region_score"].
    recorder: called once per round with (photon ids, starts, ends, energies, events, materials),
    one segment per photon still alive, as the reference's loop records them
    (chatcarlo/transport.py:219-229); events index EVENTS."""
    if grid is not None and not exact_tally and tally_rng is None:
        tally_rng = rng.spawn(1)[0]
    n = pos.shape[0]
    alive = np.ones(n, dtype=bool)
    tau = -np.log(rng.random(n))
    n_scatter = np.zeros(n, dtype=np.int64)
    absorbed = np.zeros(n, dtype=bool)
    escaped = np.zeros(n, dtype=bool)
    edep = {}
    n_fl = 0
    n_mat = len(pack.materials)
    rounds = 0
    n_seg = n_vox = 0      # flight segments, and (segment, voxel) pieces of the exact tally
    rscore = np.zeros(n)
    if region is not None:
        rc, rr = np.asarray(region[0], dtype=float), float(region[1])
        inv_vol = 1.0 / (4.0 / 3.0 * np.pi * rr ** 3)
    while np.any(alive):
        rounds += 1
        idx = np.where(alive)[0]
        n_seg += len(idx)
        o, d, e = pos[idx], dirv[idx], energy[idx]
        mat = geom.material_at(o)
        xs = core.xs_lookup(pack, mat, e)
        mu = xs["mu"]
        t_b, esc = geom.next_boundary(o, d)
        mu_safe = np.where(mu > 0, mu, 1e-30)
        tau_b = mu * t_b
        inter = tau[idx] < tau_b
        ds = np.where(inter, tau[idx] / mu_safe, t_b)
        ends = o + d * ds[:, None]
        if region is not None:
            rscore[idx] += core.h10_coeff(pack, e) * core.sphere_chord(o, d, ds, rc, rr) * inv_vol
        if grid is not None:
            wk = e * xs["mu_en"]
            wh = core.h10_coeff(pack, e)
            if exact_tally:
                n_vox += core.overlap_tally(grid.kerma, grid.origin, grid.voxel, o, d, ds, wk)
                core.overlap_tally(grid.h10, grid.origin, grid.voxel, o, d, ds, wh)
            else:
                core.stratified_tally(grid.kerma, grid.origin, grid.voxel, o, d, ds, wk, tally_rng)
                core.stratified_tally(grid.h10, grid.origin, grid.voxel, o, d, ds, wh, tally_rng)
        if recorder is not None:
            ev = np.where(inter, 0, np.where(esc, 1, 0))   # interaction types filled in below
            seg = (idx.copy(), o.copy(), ends.copy(), e.copy(), ev, mat.copy())
        pos[idx] = ends
        non = ~inter
        g = idx[non]
        tau[g] -= tau_b[non]
        pos[g] += dirv[g] * 1e-6
        en = idx[non & esc]
        alive[en] = False
        escaped[en] = True
        ii = idx[inter]
        if len(ii) == 0:
            if recorder is not None:
                recorder(*seg)
            continue
        mat_i, e_i = mat[inter], e[inter]
        r = rng.random(len(ii))
        typ = core.interaction_type(xs["p_photo"][inter], xs["p_compt"][inter], r)
        if recorder is not None:
            seg[4][inter] = np.where(typ == 0, 2, np.where(typ == 1, 4, 5))
        ph = typ == 0
        if np.any(ph):
            pidx = ii[ph]
            if fluorescence:
                _, emit, eline = core.fluor_sample(pack, mat_i[ph], e_i[ph], _draw_gen(rng))
            else:
                emit = np.zeros(len(pidx), dtype=bool)
                eline = np.zeros(len(pidx))
            dep = np.where(emit, e_i[ph] - eline, e_i[ph])
            for m in np.unique(mat_i[ph]):
                edep[int(m)] = edep.get(int(m), 0.0) + float(dep[mat_i[ph] == m].sum())
            n_scatter[pidx] += 1
            alive[pidx[~emit]] = False
            absorbed[pidx[~emit]] = True
            if recorder is not None:
                pos_in = np.flatnonzero(inter)[ph]
                seg[4][pos_in[emit]] = 3
            ei = pidx[emit]
            if len(ei):
                n_fl += len(ei)
                energy[ei] = eline[emit]
                dirv[ei] = core.isotropic_u(rng.random(len(ei)), rng.random(len(ei)))
                tau[ei] = -np.log(rng.random(len(ei)))
        co = typ == 1
        if np.any(co):
            cidx = ii[co]
            ec = e_i[co]
            eps, cos, _, _ = core.compton_sample(pack, mat_i[co], ec, _draw_gen(rng), k_trials)
            enew = ec * eps
            for m in np.unique(mat_i[co]):
                sel = mat_i[co] == m
                edep[int(m)] = edep.get(int(m), 0.0) + float((ec - enew)[sel].sum())
            dirv[cidx] = core.rotate_u(dirv[cidx], cos, rng.random(len(cidx)))
            energy[cidx] = enew
            tau[cidx] = -np.log(rng.random(len(cidx)))
            n_scatter[cidx] += 1
        ra = typ == 2
        if np.any(ra):
            ridx = ii[ra]
            cos, _, _ = core.rayleigh_sample(pack, mat_i[ra], e_i[ra], _draw_gen(rng), k_trials)
            dirv[ridx] = core.rotate_u(dirv[ridx], cos, rng.random(len(ridx)))
            tau[ridx] = -np.log(rng.random(len(ridx)))
            n_scatter[ridx] += 1
        if recorder is not None:
            recorder(*seg)
    return dict(n_scatter=n_scatter, absorbed=absorbed, escaped=escaped, final_energy=energy,
                energy_deposited=edep, n_fluorescence=n_fl, rounds=rounds, n_mat=n_mat,
                region_score=rscore, n_segments=n_seg, n_voxel_pieces=n_vox)


def run(src, geom, pack, spec_e, spec_cdf, n_histories, seed, batch_size=200_000, grid=None,
        fluorescence=True, exact_tally=True, region=None):
    """chatcarlo/transport.py:257-312 + the serial branch of :358-474."""
    rng = np.random.default_rng(seed)
    edep = {}
    sums, sum2 = {}, {}
    n_abs = n_esc = scat = n_fl = n_seg = n_vox = 0
    n_b = 0
    r_sum = r_sum2 = 0.0
    remaining = n_histories
    while remaining > 0:
        nb = min(batch_size, remaining)
        remaining -= nb
        pos, dirv, en = sample_source(src, nb, rng, spec_e, spec_cdf)
        r = transport_photons(pos, dirv, en, geom, pack, rng, grid=grid,
                              fluorescence=fluorescence, exact_tally=exact_tally, region=region)
        r_sum += float(r["region_score"].sum())
        r_sum2 += float((r["region_score"] ** 2).sum())
        for k, v in r["energy_deposited"].items():
            edep[k] = edep.get(k, 0.0) + v
        for k in set(sums) | set(r["energy_deposited"]):
            s = r["energy_deposited"].get(k, 0.0)
            sums[k] = sums.get(k, 0.0) + s
            sum2[k] = sum2.get(k, 0.0) + s * s / nb
        n_b += 1
        if grid is not None:
            grid.end_batch(nb)
        n_fl += r["n_fluorescence"]
        n_seg += r["n_segments"]
        n_vox += r["n_voxel_pieces"]
        n_abs += int(r["absorbed"].sum())
        n_esc += int(r["escaped"].sum())
        scat += int(r["n_scatter"].sum())
    return dict(energy_deposited_keV=edep, sums=sums, sum2=sum2, n_batches=n_b,
                n_absorbed=n_abs, n_escaped=n_esc, scatter_sum=scat, n_fluorescence=n_fl,
                n_histories=n_histories, region_sum=r_sum, region_sum2=r_sum2,
                n_segments=n_seg, n_voxel_pieces=n_vox)
