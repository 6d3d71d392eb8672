"""Pins the oracle's ray geometry (intersect_*, OGeom.material_at, OGeom.next_boundary):
hand-computed roots of single primitives, and next_boundary against a brute-force march of
random rays through a nested scene, where a boundary is wherever the set of primitives holding
the point changes, that set computed here from the primitives' defining inequalities."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import core

RNG = np.random.default_rng(4242)
X, Y, Z = 0, 1, 2


def _box(c, s):
    c, s = np.asarray(c, float), np.asarray(s, float)
    return (core.BOX, 0, 1, np.concatenate([c - s / 2, c + s / 2]))


def _sphere(c, r, m=2):
    return (core.SPHERE, 0, m, np.array([*c, r], float))


def _cyl(c, r, h, ax, m=3):
    return (core.CYLINDER, ax, m, np.array([*c, r, h / 2], float))


def test_single_primitive_roots():
    """Entry and exit distances worked out by hand."""
    o = np.array([[-15.0, 0.0, 0.0]])
    ex = np.array([[1.0, 0.0, 0.0]])
    t1, t2 = core.intersect_box(o, ex, np.full(3, -5.0), np.full(3, 5.0))
    assert (t1[0], t2[0]) == (10.0, 20.0)
    t1, t2 = core.intersect_sphere(np.array([[-10.0, 0.0, 0.0]]), ex, np.zeros(3), 2.0)
    assert np.allclose([t1[0], t2[0]], [8.0, 12.0], atol=1e-12)
    # sphere off-centre: impact parameter 1.2 on r = 2 gives a half chord of 1.6
    t1, t2 = core.intersect_sphere(np.array([[-10.0, 1.2, 0.0]]), ex, np.zeros(3), 2.0)
    assert np.allclose([t1[0], t2[0]], [8.4, 11.6], atol=1e-12)
    # z cylinder r = 3, half height 4: through the side along x, through the caps along z
    t1, t2 = core.intersect_cylinder(np.array([[-10.0, 0.0, 0.0]]), ex, np.zeros(3), 3.0, 4.0, Z)
    assert np.allclose([t1[0], t2[0]], [7.0, 13.0], atol=1e-12)
    t1, t2 = core.intersect_cylinder(np.array([[0.0, 0.0, -10.0]]), np.array([[0.0, 0.0, 1.0]]), np.zeros(3),
                                     3.0, 4.0, Z)
    assert np.allclose([t1[0], t2[0]], [6.0, 14.0], atol=1e-12)
    # 3-4-5 diagonal through an x cylinder's side
    d = np.array([[0.0, 0.6, 0.8]])
    t1, t2 = core.intersect_cylinder(np.array([[0.0, -6.0, -8.0]]), d, np.zeros(3), 5.0, 10.0, X)
    assert np.allclose([t1[0], t2[0]], [5.0, 15.0], atol=1e-12)
    # parallel to the axis outside the radius, and pointing away: no root ahead
    t1, t2 = core.intersect_cylinder(np.array([[0.0, 4.0, -10.0]]), np.array([[0.0, 0.0, 1.0]]), np.zeros(3),
                                     3.0, 4.0, Z)
    assert np.isnan(t1[0]) or not (t2[0] > 0 and t1[0] <= t2[0])


SCENE = [_box([0, 0, 0], [60, 50, 40]),
         _sphere([10, 5, 0], 8.0),
         _sphere([10, 5, 0], 3.0, m=4),
         _cyl([-15, 0, 0], 4.0, 30.0, Z),
         _cyl([0, -12, 5], 3.0, 40.0, X, m=5),
         _cyl([-20, 10, -5], 5.0, 12.0, Y, m=6)]
WORLD = (np.full(3, -50.0), np.full(3, 50.0))


def _inside(prim, pts):
    kind, ax, _, p = prim
    if kind == core.BOX:
        return np.all((pts >= p[:3]) & (pts <= p[3:]), axis=1)
    if kind == core.SPHERE:
        return np.linalg.norm(pts - p[:3], axis=1) <= p[3]
    o = [k for k in range(3) if k != ax]
    return (np.hypot(pts[:, o[0]] - p[o[0]], pts[:, o[1]] - p[o[1]]) <= p[3]) & (np.abs(pts[:, ax] - p[ax]) <= p[4])


def _signature(pts):
    return np.stack([_inside(pr, pts) for pr in SCENE], axis=1)


def _first_change(o, d, t_end, step=0.05):
    """First t in (0, t_end) where the containment set changes, by a march and a bisection."""
    t = np.arange(step, t_end, step)
    sig = _signature(o + t[:, None] * d)
    s0 = _signature((o + 1e-9 * d)[None, :])[0]
    ch = np.where(np.any(sig != s0, axis=1))[0]
    if len(ch) == 0:
        return np.inf
    lo, hi = (t[ch[0] - 1] if ch[0] > 0 else 0.0), t[ch[0]]
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if np.any(_signature((o + mid * d)[None, :])[0] != s0):
            hi = mid
        else:
            lo = mid
    return hi


def test_next_boundary_against_a_brute_force_march():
    g = core.OGeom(SCENE, *WORLD, background=0)
    n = 400
    o = RNG.uniform(-35.0, 35.0, (n, 3))
    d = RNG.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t_ref, esc = g.next_boundary(o, d, eps=1e-6)
    # the world exit on its own, from the slab formula
    with np.errstate(divide="ignore"):
        t_world = np.min(np.maximum((WORLD[0] - o) / d, (WORLD[1] - o) / d), axis=1)
    agree = 0
    for k in range(n):
        t_bf = _first_change(o[k], d[k], t_world[k])
        expect = min(t_bf, t_world[k])
        # the oracle never steps over a boundary the march sees
        assert t_ref[k] <= expect + 1e-7, (k, t_ref[k], expect)
        if abs(t_ref[k] - expect) <= 1e-7:
            agree += 1
            assert esc[k] == (t_bf > t_world[k])
    # the rest: a chord shorter than the march step (a grazing touch) that the march skipped
    assert agree >= n - 2, agree


def test_material_at_last_containing_primitive_wins():
    g = core.OGeom(SCENE, *WORLD, background=0)
    pts = np.array([[10.0, 5.0, 0.0],    # inner sphere inside the outer one inside the box
                    [10.0, 5.0, 5.0],    # outer sphere only
                    [-15.0, 0.0, 0.0],   # z cylinder in the box
                    [25.0, 20.0, 15.0],  # box only
                    [-45.0, 0.0, 0.0],   # background
                    [-20.0, 10.0, -5.0]])  # y cylinder
    assert g.material_at(pts).tolist() == [4, 2, 3, 1, 0, 6]
    pts = RNG.uniform(-50.0, 50.0, (20000, 3))
    sig = _signature(pts)
    mats = np.array([pr[2] for pr in SCENE])
    last = np.where(sig.any(axis=1), mats[sig.shape[1] - 1 - np.argmax(sig[:, ::-1], axis=1)], 0)
    np.testing.assert_array_equal(g.material_at(pts), last)
