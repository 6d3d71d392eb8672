"""Oracle core: geometry, cross-section lookup, samplers, tallies, statistics (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinning of each function (tests/test_oracle_golden.py, tests/test_oracle_geometry.py,
tests/test_cpu_host.py, tests/test_oracle_mlp.py):
* intersect_* / material_at / next_boundary: hand-computed roots of single
  primitives, and next_boundary against a brute-force march of random rays through a
  nested scene (boundaries where the set of containing primitives changes), material_at
  against last-wins containment from the defining inequalities (tests/test_oracle_geometry.py).
* xs_lookup: exact at the table nodes (equals the pack node values), agrees with
  the pack builder's independent log-log interpolation between nodes, matches the
  NIST spot values of the reference's test_materials.py:13-39 and the reference's
  logged linear mu golden values (docs/egs5_crosscheck/*.log, *.json).
* compton_u / rayleigh_u / fluor_u: sampled distributions against closed forms
  (Klein-Nishina with S=Z, inverse-CDF of F^2 by quadrature, K-line shares),
  the bound-Compton mean transfer anchor t_ref = 0.095928 of the reference's
  tests/test_bound_compton.py:66-80 (within the data-model tolerance).
* rotate_u / isotropic_u: unit norm, cos(d, d') = cos theta exactly, isotropy.
* overlap_tally: total conservation, analytic single-voxel chords, agreement
  with the reference's stratified estimator (chatcarlo/tally.py:248-293) in
  expectation.
* batch statistics: brute force at batch size 1 (reference tests/test_uncertainty.py:22-50).
* sphere_chord: hand-computed chords and Cauchy's mean chord 4r/3 for isotropic
  uniform lines (tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import math

import numpy as np

HC = 12.3984193     # chatcarlo/physics.py:35
MEC2 = 511.0        # chatcarlo/physics.py:34
FLUOR_CUTOFF = 5.0  # chatcarlo/physics.py:38

BOX, SPHERE, CYLINDER = 0, 1, 2


# ---------------------------------------------------------------------------
# geometry (chatcarlo/geometry.py)
# ---------------------------------------------------------------------------
def intersect_box(o, d, lo, hi, eps=1e-12):
    """Slab method, chatcarlo/geometry.py:20-42 (bounds given as lo/hi)."""
    n = o.shape[0]
    t_enter = np.full(n, -np.inf)
    t_exit = np.full(n, np.inf)
    miss = np.zeros(n, dtype=bool)
    for k in range(3):
        dk, ok = d[:, k], o[:, k]
        par = np.abs(dk) < eps
        miss |= par & ((ok < lo[k]) | (ok > hi[k]))
        ds = np.where(par, 1.0, dk)
        ta = (lo[k] - ok) / ds
        tb = (hi[k] - ok) / ds
        t_enter = np.maximum(t_enter, np.where(par, -np.inf, np.minimum(ta, tb)))
        t_exit = np.minimum(t_exit, np.where(par, np.inf, np.maximum(ta, tb)))
    hit = (~miss) & (t_enter <= t_exit)
    return np.where(hit, t_enter, np.nan), np.where(hit, t_exit, np.nan)


def intersect_sphere(o, d, c, r):
    """chatcarlo/geometry.py:45-54."""
    oc = o - np.asarray(c, dtype=float)
    b = np.einsum("ij,ij->i", oc, d)
    cc = np.einsum("ij,ij->i", oc, oc) - r * r
    disc = b * b - cc
    hit = disc >= 0
    sq = np.sqrt(np.where(hit, disc, 0.0))
    t1, t2 = -b - sq, -b + sq
    return np.where(hit, np.minimum(t1, t2), np.nan), np.where(hit, np.maximum(t1, t2), np.nan)


def intersect_cylinder(o, d, c, r, half_h, axis, eps=1e-12):
    """chatcarlo/geometry.py:57-99 (height given as half height)."""
    oi = [k for k in range(3) if k != axis]
    c = np.asarray(c, dtype=float)
    ox, oy = o[:, oi[0]] - c[oi[0]], o[:, oi[1]] - c[oi[1]]
    dx, dy = d[:, oi[0]], d[:, oi[1]]
    a = dx * dx + dy * dy
    b = ox * dx + oy * dy
    cc = ox * ox + oy * oy - r * r
    nz = a < eps
    disc = b * b - a * cc
    hit_r = (~nz) & (disc >= 0)
    sq = np.sqrt(np.where(hit_r, disc, 0.0))
    a_s = np.where(nz, 1.0, a)
    t1, t2 = (-b - sq) / a_s, (-b + sq) / a_s
    ten_r = np.where(hit_r, np.minimum(t1, t2), -np.inf)
    tex_r = np.where(hit_r, np.maximum(t1, t2), np.inf)
    inside_line = nz & (cc <= 0)
    ten_r = np.where(inside_line, -np.inf, ten_r)
    tex_r = np.where(inside_line, np.inf, tex_r)
    miss_r = (~hit_r) & (~inside_line)
    lo_h, hi_h = c[axis] - half_h, c[axis] + half_h
    oz, dz = o[:, axis], d[:, axis]
    par = np.abs(dz) < eps
    miss_z = par & ((oz < lo_h) | (oz > hi_h))
    dzs = np.where(par, 1.0, dz)
    ta, tb = (lo_h - oz) / dzs, (hi_h - oz) / dzs
    ten_z = np.where(par, -np.inf, np.minimum(ta, tb))
    tex_z = np.where(par, np.inf, np.maximum(ta, tb))
    ten, tex = np.maximum(ten_r, ten_z), np.minimum(tex_r, tex_z)
    miss = miss_r | miss_z | (ten > tex)
    return np.where(miss, np.nan, ten), np.where(miss, np.nan, tex)


class OGeom:
    """Primitive list with last-wins containment (chatcarlo/geometry.py:121-231).

    prims: iterable of (kind, axis, material_id, params) where params are
    box (lo3, hi3), sphere (c3, r), cylinder (c3, r, half_h).
    world_lo/hi: world box; background: material id outside every primitive.
    """

    def __init__(self, prims, world_lo, world_hi, background):
        self.prims = [(int(k), int(ax), int(m), np.asarray(p, dtype=float)) for k, ax, m, p in prims]
        self.world_lo = np.asarray(world_lo, dtype=float)
        self.world_hi = np.asarray(world_hi, dtype=float)
        self.background = int(background)

    @classmethod
    def from_packed(cls, prim_arr, world):
        """From the product's packed float32 arrays (device parameters, upcast)."""
        prims = []
        for rec in prim_arr:
            kind, axis, mat = int(rec["kind"]), int(rec["axis"]), int(rec["material"])
            p = rec["p"].astype(np.float64)
            prims.append((kind, axis, mat, p))
        return cls(prims, world["lo"].astype(np.float64), world["hi"].astype(np.float64),
                   int(world["background"]))

    def _contains(self, prim, pts):
        kind, ax, _, p = prim
        if kind == BOX:
            return np.all((pts >= p[0:3]) & (pts <= p[3:6]), axis=1)
        if kind == SPHERE:
            return np.sum((pts - p[0:3]) ** 2, axis=1) <= p[3] * p[3]
        oi = [k for k in range(3) if k != ax]
        r2 = (pts[:, oi[0]] - p[oi[0]]) ** 2 + (pts[:, oi[1]] - p[oi[1]]) ** 2
        return (r2 <= p[3] * p[3]) & (np.abs(pts[:, ax] - p[ax]) <= p[4])

    def _intersect(self, prim, o, d):
        kind, ax, _, p = prim
        if kind == BOX:
            return intersect_box(o, d, p[0:3], p[3:6])
        if kind == SPHERE:
            return intersect_sphere(o, d, p[0:3], p[3])
        return intersect_cylinder(o, d, p[0:3], p[3], p[4], ax)

    def material_at(self, pts):
        """chatcarlo/geometry.py:205-212: last primitive containing the point wins."""
        mat = np.full(pts.shape[0], self.background, dtype=np.int64)
        for prim in self.prims:
            m = self._contains(prim, pts)
            mat[m] = prim[2]
        return mat

    def next_boundary(self, o, d, eps=1e-6):
        """chatcarlo/geometry.py:214-231: min positive root (> eps) vs world exit."""
        n = o.shape[0]
        t_obj = np.full(n, np.inf)
        for prim in self.prims:
            t1, t2 = self._intersect(prim, o, d)
            for t in (t1, t2):
                valid = ~np.isnan(t) & (t > eps)
                t_obj = np.where(valid & (t < t_obj), t, t_obj)
        _, t_exit = intersect_box(o, d, self.world_lo, self.world_hi)
        t_exit = np.where(np.isnan(t_exit), np.inf, t_exit)
        escape = t_exit <= t_obj
        return np.minimum(t_obj, t_exit), escape


# ---------------------------------------------------------------------------
# cross-section lookup (chatcarlo/materials.py:228-262, device table form)
# ---------------------------------------------------------------------------
PHOTO, COMPT, RAYL, MUEN, TOTAL = 0, 1, 2, 3, 4
_EDGE_PART = {PHOTO: 0, MUEN: 1, TOTAL: 2}


def energy_interval(pack, e):
    """Interval index i with e_node[i] <= E < e_node[i+1] and f = ln(E/E_i)/ln(E_i+1/E_i).

    Raises ValueError outside [1, 150] keV like chatcarlo/materials.py:234-237.
    """
    e = np.asarray(e, dtype=float)
    if e.size and (e.min() < pack.e_node[0] or e.max() > pack.e_node[-1]):
        raise ValueError(f"energy {e.min():.3g}-{e.max():.3g} keV outside the cross-section "
                         f"table range [{pack.e_node[0]:.3g}, {pack.e_node[-1]:.3g}] keV")
    n = len(pack.e_node)
    i = np.clip(np.searchsorted(pack.e_node, e, side="right") - 1, 0, n - 2)
    f = np.log(e / pack.e_node[i]) / pack.dl_node[i]
    return i, f


def xs_part(pack, mat, e, part, i=None, f=None):
    """One part (photo/compt/rayl/muen/total) in cm^2/g at (mat, E), log-log linear."""
    mat = np.asarray(mat, dtype=np.int64)
    e = np.asarray(e, dtype=float)
    if i is None:
        i, f = energy_interval(pack, e)
    s0 = pack.sig[mat, part, i]
    val = s0 * np.exp(f * pack.slope[mat, part, i])
    if part in _EDGE_PART:
        k = pack.edge_of_interval[mat, i]
        has = k >= 0
        if np.any(has):
            p = _EDGE_PART[part]
            mm, kk, ii, ee = mat[has], k[has], i[has], e[has]
            ledge = pack.edge_e[mm, kk]
            below = ee < ledge
            fb = np.log(ee / pack.e_node[ii]) / pack.edge_dl[mm, kk, 0]
            fa = np.log(ee / ledge) / pack.edge_dl[mm, kk, 1]
            vb = pack.sig[mm, part, ii] * np.exp(fb * pack.edge_slope[mm, kk, p, 0])
            va = pack.edge_sig[mm, kk, p, 1] * np.exp(fa * pack.edge_slope[mm, kk, p, 1])
            val = val.copy()
            val[has] = np.where(below, vb, va)
    return val


def xs_lookup(pack, mat, e):
    """mu [1/cm], p_photo, p_compt, parts (photo, compt, rayl) [cm^2/g], mu_en [1/cm], mu_total [1/cm].

    Interaction probabilities follow chatcarlo/transport.py:72-85 (parts / sum).
    """
    mat = np.asarray(mat, dtype=np.int64)
    i, f = energy_interval(pack, e)
    ph = xs_part(pack, mat, e, PHOTO, i, f)
    co = xs_part(pack, mat, e, COMPT, i, f)
    ra = xs_part(pack, mat, e, RAYL, i, f)
    tot = ph + co + ra
    rho = pack.density[mat]
    mu_parts = tot * rho
    mu_total = xs_part(pack, mat, e, TOTAL, i, f) * rho
    muen = xs_part(pack, mat, e, MUEN, i, f) * rho
    ts = np.where(tot > 0, tot, 1.0)
    return dict(mu=mu_parts, p_photo=ph / ts, p_compt=co / ts, photo=ph, compt=co, rayl=ra,
                mu_en=muen, mu_total=mu_total)


def h10_coeff(pack, e):
    """h*(10)/Phi [pSv cm^2], chatcarlo/dose_coefficients.py:32-42 (log-log, clamped)."""
    e = np.atleast_1d(np.asarray(e, dtype=float))
    ec = np.clip(e, pack.h10_e[0], pack.h10_e[-1])
    return np.exp(np.interp(np.log(ec), np.log(pack.h10_e), np.log(pack.h10_v)))


def h10_coeff_dev(pack, e):
    """Same function in the device's per-interval form h_j exp(ln(E/E_j)/d_j * s_j)."""
    e = np.atleast_1d(np.asarray(e, dtype=float))
    ec = np.clip(e, pack.h10_e[0], pack.h10_e[-1])
    j = np.clip(np.searchsorted(pack.h10_e, ec, side="right") - 1, 0, len(pack.h10_e) - 2)
    f = np.log(ec / pack.h10_e[j]) / pack.h10_dl[j]
    return pack.h10_v[j] * np.exp(f * pack.h10_slope[j])


# ---------------------------------------------------------------------------
# samplers fed explicit uniforms (chatcarlo/physics.py)
# ---------------------------------------------------------------------------
def sel_cdf_at(pack, mat, kind, i, f):
    """Cumulative element-selection fractions at E (linear in f between nodes)."""
    cdf = pack.sel_cdf[mat]
    return cdf[kind, i] + f * (cdf[kind, i + 1] - cdf[kind, i])


def select_element(cdf_row, r):
    """idx = #(cdf < r), clipped (chatcarlo/physics.py:88-91)."""
    return min(int(np.sum(r > cdf_row)), len(cdf_row) - 1)


def select_elements(pack, mat, kind, i, f, r):
    """Vectorised element pick per photon; returns global element ids."""
    mat = np.asarray(mat, dtype=np.int64)
    out = np.zeros(len(mat), dtype=np.int64)
    for m in np.unique(mat):
        sel = np.where(mat == m)[0]
        cdf = pack.sel_cdf[m]
        rows = cdf[kind, i[sel]] + f[sel][:, None] * (cdf[kind, i[sel] + 1] - cdf[kind, i[sel]])
        k = np.minimum(np.sum(r[sel][:, None] > rows, axis=1), rows.shape[1] - 1)
        out[sel] = np.asarray(pack.mat_elem[m], dtype=np.int64)[k]
    return out


def _matrix_draw(u):
    return lambda idx, slot: u[idx, slot]


def s_over_z(pack, el, q):
    """np.interp of S(Z,q)/Z on the q table, per element (chatcarlo/physics.py:124-127)."""
    out = np.empty(len(q))
    for z in np.unique(el):
        m = el == z
        out[m] = np.interp(q[m], pack.q_s, pack.sq_over_z[z])
    return out


def compton_sample(pack, mat, e, draw, k_trials):
    """Bound Compton (chatcarlo/physics.py:41-56, 78-135) with a uniform source.

    Slot 0 picks the element; trial t uses slots 1+3t, 2+3t, 3+3t as
    (xi1, xi2, xi3).  Returns eps, cos, element id, trials (0 = cap hit).
    """
    e = np.asarray(e, dtype=float)
    n = len(e)
    i, f = energy_interval(pack, e)
    el = select_elements(pack, mat, COMPT, i, f, draw(np.arange(n), 0))
    a = e / MEC2
    emin = 1.0 / (1.0 + 2.0 * a)
    env = 1.0 / emin + emin
    eps_o, cos_o = np.zeros(n), np.zeros(n)
    tr_o = np.zeros(n, dtype=np.int64)
    pending = np.arange(n)
    for t in range(k_trials):
        if len(pending) == 0:
            break
        x1 = draw(pending, 1 + 3 * t)
        x2 = draw(pending, 2 + 3 * t)
        x3 = draw(pending, 3 + 3 * t)
        ap, em = a[pending], emin[pending]
        eps = em + x1 * (1.0 - em)
        omc = (1.0 / eps - 1.0) / ap
        cos = 1.0 - omc
        g = 1.0 / eps + eps - (1.0 - cos * cos)
        acc_kn = x2 * env[pending] <= g
        q = e[pending] * np.sqrt(np.clip(omc, 0.0, 2.0) / 2.0) / HC
        acc = acc_kn & (x3 <= s_over_z(pack, el[pending], q))
        done = pending[acc]
        eps_o[done], cos_o[done], tr_o[done] = eps[acc], cos[acc], t + 1
        pending = pending[~acc]
    return eps_o, cos_o, el, tr_o


def compton_u(pack, mat, e, u, k_trials):
    return compton_sample(pack, mat, e, _matrix_draw(np.asarray(u)), k_trials)


def rayleigh_sample(pack, mat, e, draw, k_trials):
    """Two-stage Rayleigh (chatcarlo/physics.py:223-236, 281-351) with a uniform source.

    Slot 0 picks the element; trial t uses slot 1+2t (inverse A(x)) and slot
    2+2t (angular factor (1+c^2)/2).  Raises ValueError if x_max exceeds the table.
    """
    e = np.asarray(e, dtype=float)
    n = len(e)
    i, f = energy_interval(pack, e)
    el = select_elements(pack, mat, RAYL, i, f, draw(np.arange(n), 0))
    xg = pack.ray_x
    x_max = (e / HC) ** 2
    if n and x_max.max() > xg[-1]:
        raise ValueError("q(E) beyond the Rayleigh form-factor table")
    a_cut = np.empty(n)
    for z in np.unique(el):
        m = el == z
        a_cut[m] = np.interp(x_max[m], xg, pack.ray_a[z])
    cos_o = np.zeros(n)
    tr_o = np.zeros(n, dtype=np.int64)
    pending = np.arange(n)
    for t in range(k_trials):
        if len(pending) == 0:
            break
        x1 = draw(pending, 1 + 2 * t)
        x2 = draw(pending, 2 + 2 * t)
        x = np.empty(len(pending))
        elp = el[pending]
        for z in np.unique(elp):
            m = elp == z
            x[m] = np.interp(x1[m] * a_cut[pending][m], pack.ray_a[z], xg)
        xm = x_max[pending]
        x = np.minimum(x, xm)
        c = np.clip(1.0 - 2.0 * x / xm, -1.0, 1.0)
        acc = x2 <= (1.0 + c * c) / 2.0
        done = pending[acc]
        cos_o[done], tr_o[done] = c[acc], t + 1
        pending = pending[~acc]
    return cos_o, el, tr_o


def rayleigh_u(pack, mat, e, u, k_trials):
    return rayleigh_sample(pack, mat, e, _matrix_draw(np.asarray(u)), k_trials)


def k_share(pack, el, e):
    """K-shell share of photoabsorption by element el at energy e above its K edge: the pack's
    nodes, spaced 1/fl_kinv in ln E from ln(edge), linear in ln E between them (the pack's
    restatement of CS_Photo_Partial(Z,K,E)/CS_Photo(Z,E), chatcarlo/physics.py:193-197).
    Zero where the element has no fluorescence data."""
    el = np.asarray(el)
    e = np.asarray(e, dtype=float)
    edge = pack.fl_edge[el]
    ok = edge > 0
    ln_edge = np.log(np.where(ok, edge, 1.0))
    if pack.dtype == "f32":   # the device keeps ln(edge) as a float32 table value
        ln_edge = ln_edge.astype(np.float32).astype(np.float64)
    t = (np.log(e) - ln_edge) * pack.fl_kinv[el]
    nk = pack.fl_kfrac.shape[1]
    j = np.clip(np.floor(t).astype(int), 0, nk - 2)
    k0 = pack.fl_kfrac[el, j]
    k1 = pack.fl_kfrac[el, j + 1]
    return np.where(ok, k0 + (t - j) * (k1 - k0), 0.0)


def fluor_sample(pack, mat, e, draw):
    """K fluorescence after photoabsorption (chatcarlo/physics.py:138-212).

    Slot 0 picks the absorbing element, slot 1 the K-shell test (k_frac),
    slot 2 the radiative test (omega_K), slot 3 the line.  Returns element,
    emit, E_line.
    """
    e = np.asarray(e, dtype=float)
    n = len(e)
    i, f = energy_interval(pack, e)
    allidx = np.arange(n)
    el = select_elements(pack, mat, PHOTO, i, f, draw(allidx, 0))
    u1, u2, u3 = draw(allidx, 1), draw(allidx, 2), draw(allidx, 3)
    edge = pack.fl_edge[el]
    cand = (edge > 0) & (e > edge) & (u1 < k_share(pack, el, e)) & (u2 < pack.fl_omega[el])
    emit = np.zeros(n, dtype=bool)
    eline = np.zeros(n)
    for j in np.where(cand)[0]:
        z = el[j]
        li = select_element(pack.fl_cdf[z], u3[j])
        ln = pack.fl_lines[z, li]
        if ln >= FLUOR_CUTOFF:
            emit[j] = True
            eline[j] = ln
    return el, emit, eline


def fluor_u(pack, mat, e, u):
    return fluor_sample(pack, mat, e, _matrix_draw(np.asarray(u)))


def rotate_u(d, cos_t, u_phi):
    """New direction from (cos theta, phi = 2 pi u), chatcarlo/physics.py:354-367."""
    d = np.asarray(d, dtype=float)
    sin_t = np.sqrt(np.clip(1.0 - cos_t ** 2, 0.0, None))
    phi = 2.0 * np.pi * u_phi
    up = np.where((np.abs(d[:, 2]) < 0.999)[:, None], np.array([0.0, 0.0, 1.0]),
                  np.array([1.0, 0.0, 0.0]))
    uu = np.cross(up, d)
    uu /= np.linalg.norm(uu, axis=1, keepdims=True)
    vv = np.cross(d, uu)
    nd = ((sin_t * np.cos(phi))[:, None] * uu + (sin_t * np.sin(phi))[:, None] * vv
          + cos_t[:, None] * d)
    return nd / np.linalg.norm(nd, axis=1, keepdims=True)


def isotropic_u(u1, u2):
    """chatcarlo/physics.py:215-220: cos = U(-1,1), phi = U(0, 2 pi)."""
    c = -1.0 + 2.0 * np.asarray(u1, dtype=float)
    s = np.sqrt(np.clip(1.0 - c * c, 0.0, None))
    phi = 2.0 * np.pi * np.asarray(u2, dtype=float)
    return np.column_stack([s * np.cos(phi), s * np.sin(phi), c])


def interaction_type(p_photo, p_compt, r):
    """0 photo, 1 compton, 2 rayleigh (chatcarlo/transport.py:165-170)."""
    return np.where(r < p_photo, 0, np.where(r < p_photo + p_compt, 1, 2))


# ---------------------------------------------------------------------------
# tallies (chatcarlo/tally.py)
# ---------------------------------------------------------------------------
def overlap_tally(target, origin_cm, voxel_cm, o, d, length, weight):
    """Exact segment-voxel overlap scoring: target[v] += weight * overlap(v).
    Returns the number of (segment, voxel) pieces of positive length.

    The expectation of the reference's stratified estimator
    (chatcarlo/tally.py:248-293) is exactly this overlap; the exact form is the
    one docs/egs5_crosscheck/run_chatcarlo_pdd60.py:49-73 uses.
    """
    o = np.asarray(o, dtype=float)
    d = np.asarray(d, dtype=float)
    length = np.asarray(length, dtype=float)
    weight = np.asarray(weight, dtype=float)
    n = o.shape[0]
    if n == 0:
        return 0
    shape = np.array(target.shape)
    org = np.asarray(origin_cm, dtype=float)
    hi = org + shape * voxel_cm
    # clip each segment to the grid box
    t0 = np.zeros(n)
    t1 = length.copy()
    for k in range(3):
        dk = d[:, k]
        par = np.abs(dk) < 1e-300
        out = par & ((o[:, k] < org[k]) | (o[:, k] > hi[k]))
        t1 = np.where(out, -1.0, t1)
        ds = np.where(par, 1.0, dk)
        ta = (org[k] - o[:, k]) / ds
        tb = (hi[k] - o[:, k]) / ds
        t0 = np.where(par, t0, np.maximum(t0, np.minimum(ta, tb)))
        t1 = np.where(par, t1, np.minimum(t1, np.maximum(ta, tb)))
    keep = t1 > t0
    if not np.any(keep):
        return 0
    o, d, t0, t1, weight = o[keep], d[keep], t0[keep], t1[keep], weight[keep]
    n = o.shape[0]
    ts = [t0, t1]
    seg = [np.arange(n), np.arange(n)]
    for k in range(3):
        dk = d[:, k]
        nzk = np.abs(dk) > 0
        c0 = (o[:, k] + dk * t0 - org[k]) / voxel_cm
        c1 = (o[:, k] + dk * t1 - org[k]) / voxel_cm
        lo_p = np.floor(np.minimum(c0, c1)) + 1
        hi_p = np.ceil(np.maximum(c0, c1)) - 1
        cnt = np.where(nzk, np.maximum(hi_p - lo_p + 1, 0), 0).astype(np.int64)
        if cnt.sum() == 0:
            continue
        sid = np.repeat(np.arange(n), cnt)
        start = np.repeat(lo_p, cnt)
        offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        plane = org[k] + (start + offs) * voxel_cm
        t = (plane - o[sid, k]) / dk[sid]
        ts.append(t)
        seg.append(sid)
    t_all = np.concatenate(ts)
    s_all = np.concatenate(seg)
    order = np.lexsort((t_all, s_all))
    t_all, s_all = t_all[order], s_all[order]
    same = s_all[1:] == s_all[:-1]
    ta, tb, sid = t_all[:-1][same], t_all[1:][same], s_all[:-1][same]
    ln = tb - ta
    good = ln > 0
    ta, tb, sid, ln = ta[good], tb[good], sid[good], ln[good]
    mid = o[sid] + d[sid] * (0.5 * (ta + tb))[:, None]
    idx = np.floor((mid - org) / voxel_cm).astype(np.int64)
    idx = np.clip(idx, 0, shape - 1)
    flat = np.ravel_multi_index((idx[:, 0], idx[:, 1], idx[:, 2]), target.shape)
    np.add.at(target.reshape(-1), flat, weight[sid] * ln)
    return int(len(ln))


def stratified_tally(target, origin_cm, voxel_cm, o, d, length, weight, rng,
                     substep_cm=None, max_substeps=40):
    """The reference's stratified sub-step estimator, chatcarlo/tally.py:248-293."""
    n = o.shape[0]
    if n == 0:
        return
    if substep_cm is None:
        substep_cm = voxel_cm / 2.0
    nsub = np.clip(np.ceil(length / substep_cm).astype(int), 1, max_substeps)
    mx = int(nsub.max())
    j = np.arange(mx)
    frac = (j[None, :] + rng.random((n, mx))) / nsub[:, None]
    valid = j[None, :] < nsub[:, None]
    pts = o[:, None, :] + d[:, None, :] * (length[:, None] * frac)[:, :, None]
    sw = weight * (length / nsub)
    pts = pts.reshape(-1, 3)
    wf = np.broadcast_to(sw[:, None], (n, mx)).reshape(-1)
    idx = np.floor((pts - np.asarray(origin_cm)) / voxel_cm).astype(int)
    shape = np.array(target.shape)
    ok = valid.reshape(-1) & np.all((idx >= 0) & (idx < shape), axis=1)
    idx, wf = idx[ok], wf[ok]
    if len(idx) == 0:
        return
    flat = np.ravel_multi_index((idx[:, 0], idx[:, 1], idx[:, 2]), target.shape)
    np.add.at(target.reshape(-1), flat, wf)


# ---------------------------------------------------------------------------
# batch statistics (chatcarlo/tally.py:135-188)
# ---------------------------------------------------------------------------
def batch_variance(T, Q, m, n):
    T = np.asarray(T, dtype=float)
    Q = np.asarray(Q, dtype=float)
    if m < 2 or n <= 0:
        return np.full_like(T, np.nan)
    return np.clip((Q - T ** 2 / n) / (m - 1), 0.0, None)


def standard_error(T, Q, m, n):
    v = batch_variance(T, Q, m, n)
    if n <= 0:
        return np.full_like(v, np.nan)
    return np.sqrt(v / n)


def relative_error(T, Q, m, n):
    T = np.asarray(T, dtype=float)
    sem = standard_error(T, Q, m, n)
    if n <= 0:
        return np.full_like(T, np.nan)
    mean = T / n
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(mean != 0, sem / mean, np.nan)


# ---------------------------------------------------------------------------
# region tally (the operator sphere of the learned-importance workload; the
# reference scores H*(10) only on its voxel grid, chatcarlo/tally.py:248-293)
# ---------------------------------------------------------------------------
def sphere_chord(o, d, length, c, r):
    """Length of each segment [o, o + d length] inside the sphere (c, r)."""
    o = np.atleast_2d(np.asarray(o, dtype=float))
    d = np.atleast_2d(np.asarray(d, dtype=float))
    oc = o - np.asarray(c, dtype=float)
    b = np.einsum("ij,ij->i", oc, d)
    cc = np.einsum("ij,ij->i", oc, oc) - r * r
    disc = b * b - cc
    sq = np.sqrt(np.maximum(disc, 0.0))
    lo = np.maximum(-b - sq, 0.0)
    hi = np.minimum(-b + sq, np.asarray(length, dtype=float))
    return np.where((disc > 0) & (hi > lo), hi - lo, 0.0)
