"""Pins the oracle of the learned-importance layer: MLP forward/backward, Adam,
weight-window rules, source-cell weights, feature rows and the sphere chord."""
import numpy as np
import pytest

from oracle import core
from oracle import mlp as M

RNG = np.random.default_rng(5)


def _params(scale=1.0, seed=3):
    w = M.init_params(seed).astype(np.float64)
    w[M.OFF_B2:M.OFF_W3] = RNG.normal(0, 0.1, M.N_H)
    w[M.OFF_W3:M.OFF_B3] = RNG.normal(0, 0.3, M.N_H) * scale
    w[M.OFF_B3] = 0.2
    W1 = w[M.OFF_W1:M.OFF_W2].reshape(M.N_IN, M.N_H)
    W1[7] = RNG.normal(0, 0.2, M.N_H)     # first-layer bias row
    return w


def _rows(n):
    x = np.zeros((n, M.N_IN))
    x[:, :3] = RNG.uniform(-1, 1, (n, 3))
    d = RNG.normal(size=(n, 3))
    x[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    x[:, 6] = 0.5 * np.log(RNG.uniform(5, 140, n) / 50.0)
    x[:, 7] = 1.0
    return x


def test_forward_matches_the_scalar_definition():
    w = _params()
    x = _rows(6)
    f, _ = M.forward(w, x, bf16=False)
    np.testing.assert_allclose(f, M.forward_loop(w, x), rtol=1e-12, atol=1e-12)


def test_padding_rows_of_w1_do_not_act():
    w = _params()
    x = _rows(5)
    x2 = x.copy()
    x2[:, 8:] = RNG.normal(size=(5, 8))
    w2 = w.copy()
    w2[M.OFF_W1 + 8 * M.N_H:M.OFF_W2] = 7.0
    np.testing.assert_allclose(M.forward(w2, x2, False)[0], M.forward(w, x, False)[0], rtol=1e-13)


@pytest.mark.parametrize("yscale", [1.0, 3.5])
def test_fp64_gradient_matches_central_differences(yscale):
    w = _params()
    x = _rows(40)
    y = RNG.poisson(0.7, 40).astype(float)
    g = M.grad(w, x, y, yscale=yscale, gscale=0.25, bf16=False)
    h = 1e-6
    idx = np.concatenate([RNG.choice(np.arange(0, 8 * M.N_H), 25, replace=False),
                          RNG.choice(np.arange(M.OFF_W2, M.OFF_B2), 25, replace=False),
                          np.arange(M.OFF_B2, M.OFF_B2 + 10), np.arange(M.OFF_W3, M.OFF_W3 + 10), [M.OFF_B3]])
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (M.loss(wp, x, y, yscale, 0.25) - M.loss(wm, x, y, yscale, 0.25)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * (1 + abs(fd)), (k, fd, g[k])
    assert np.all(g[M.OFF_W1 + 8 * M.N_H:M.OFF_W2] == 0)


def test_bf16_gradient_is_close_to_fp64():
    w = _params()
    x = _rows(64)
    y = RNG.poisson(1.0, 64).astype(float)
    g64 = M.grad(w, x, y, bf16=False)
    g16 = M.grad(w, x, y, bf16=True)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 0.05


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.140625, 1e-3])
    got = M.bf16_round(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0   # 1 + 2^-8 sits halfway between 1 and 1 + 2^-7: the even mantissa wins
    assert got[2] == 1.0 + 2 ** -7
    assert got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[6] - 1e-3) / 1e-3 < 2 ** -8


def test_minimiser_of_the_poisson_loss_is_the_log_mean():
    # one free parameter (b3) and constant features: d/db3 sum(exp(b3) - y b3) = 0 at b3 = ln mean(y)
    y = RNG.poisson(2.3, 500).astype(float)
    w = np.zeros(M.N_PARAMS)
    x = _rows(500)
    b = np.log(y.mean())
    w[M.OFF_B3] = b
    g = M.grad(w, x, y, bf16=False)
    assert abs(g[M.OFF_B3]) < 1e-9 * len(y)


def test_adam_constant_gradient_closed_form():
    # with a constant gradient g, m_t/(1-b1^t) = g and v_t/(1-b2^t) = g^2: each step moves lr g/(|g|+eps)
    w = np.zeros(M.N_PARAMS)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    g = RNG.normal(size=M.N_PARAMS)
    g[M.OFF_W1 + 8 * M.N_H:M.OFF_W2] = 0
    for t in range(1, 6):
        w, m, v = M.adam_step(w, m, v, g, t, lr=0.01)
    expect = -5 * 0.01 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(w, expect, rtol=1e-6, atol=1e-12)


def test_weight_window_keeps_the_expected_weight():
    wt = np.full(200000, 0.5)
    u = RNG.random(200000)
    # splitting: n copies of w / n sum to w exactly
    w = np.full(200000, 3.7)
    n, wc, alive = M.window(w, wt, 16, u)
    assert np.all(n == 7) and np.allclose(n * wc, w)
    n, wc, _ = M.window(np.full(5, 40.0), np.full(5, 0.5), 16, u[:5])
    assert np.all(n == 16) and np.allclose(n * wc, 40.0)
    # roulette: E[w'] = w
    w = np.full(200000, 0.05)
    n, wc, alive = M.window(w, wt, 16, u)
    assert np.all(n == 1)
    assert abs(wc.mean() - 0.05) < 4 * 0.5 * np.sqrt(0.1 * 0.9 / 200000)
    assert set(np.unique(wc)) <= {0.0, 0.5}
    # inside the window: untouched
    n, wc, alive = M.window(np.full(3, 0.6), np.full(3, 0.5), 16, u[:3])
    assert np.all(n == 1) and np.all(wc == 0.6) and np.all(alive)


def test_target_weight_clamps():
    f = np.array([-100.0, 0.0, 3.0, 100.0])
    wt = M.target_weight(f, 2.0, 1e-3, 1e3)
    assert wt[0] == 1e3 and wt[1] == 2.0 and abs(wt[2] - 2 * np.exp(-3)) < 1e-15 and wt[3] == 1e-3


def test_source_cells_are_unbiased():
    imp = RNG.lognormal(0, 2, 512)
    for mix in (0.0, 0.1, 1.0):
        p, wgt = M.source_cells(imp, mix)
        assert abs(p.sum() - 1) < 1e-12
        # E_p[w * h] = sum_k p_k (1/K)/p_k h_k = mean h  for any cell function h
        h = RNG.random(512)
        assert abs(np.sum(p * wgt * h) - h.mean()) < 1e-12
    p, _ = M.source_cells(imp, 1.0)
    np.testing.assert_allclose(p, 1 / 512)


def test_feature_rows_by_hand():
    c = np.array([10.0, 0.0, 5.0])
    # photon 20 cm below the centre along -x, heading straight at it: first direction feature 1
    x = M.features([[-10.0, 0.0, 5.0]], [[1.0, 0.0, 0.0]], [50.0 * np.e ** 2], c, 100.0)
    assert np.allclose(x[0, :8], [-0.2, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0])
    assert np.all(x[0, 8:] == 0)
    # u = +x: e2 = z x u = +y, e3 = u x e2 = +z
    x = M.features([[-10.0, 0.0, 5.0]], [[0.0, 0.6, -0.8]], [50.0], c, 100.0)
    assert np.allclose(x[0, 3:7], [0.0, 0.6, -0.8, 0.0])
    # u along +z (polar case): e2 = u x x = +y, e3 = u x e2 = -x
    x = M.features([[10.0, 0.0, -95.0]], [[0.6, 0.0, 0.8]], [50.0], c, 100.0)
    assert np.allclose(x[0, :6], [0.0, 0.0, -1.0, 0.8, 0.0, -0.6])
    # the direction features are the direction in an orthonormal frame: norm 1
    d = RNG.normal(size=(100, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x = M.features(RNG.uniform(-50, 50, (100, 3)), d, np.full(100, 60.0), c, 100.0)
    assert np.allclose(np.linalg.norm(x[:, 3:6], axis=1), 1.0)


def test_sphere_chord_closed_forms():
    c, r = np.array([1.0, 2.0, 3.0]), 2.0
    o = np.array([[1.0, 2.0, -5.0], [1.0, 2.0, -5.0], [1.0, 3.0, -5.0], [1.0, 2.0, 3.0], [1.0, 5.0, -5.0]])
    d = np.array([[0, 0, 1.0], [0, 0, 1.0], [0, 0, 1.0], [1.0, 0, 0], [0, 0, 1.0]])
    L = np.array([100.0, 7.0, 100.0, 0.5, 100.0])
    got = core.sphere_chord(o, d, L, c, r)
    np.testing.assert_allclose(got, [4.0, 1.0, 2 * np.sqrt(3.0), 0.5, 0.0], atol=1e-12)


def test_sphere_chord_cauchy_mean():
    # isotropic uniform random lines hitting a sphere: mean chord = 4 r / 3 (Cauchy-Crofton)
    n = 400000
    r = 1.5
    d = RNG.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # uniform impact point on the disc orthogonal to d
    a = np.where(np.abs(d[:, 2:3]) < 0.9, np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]]))
    e1 = np.cross(d, a)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(d, e1)
    rad = r * np.sqrt(RNG.random(n))
    phi = 2 * np.pi * RNG.random(n)
    o = (rad * np.cos(phi))[:, None] * e1 + (rad * np.sin(phi))[:, None] * e2 - 10 * d
    ch = core.sphere_chord(o, d, np.full(n, 100.0), np.zeros(3), r)
    assert abs(ch.mean() - 4 * r / 3) < 4 * ch.std() / np.sqrt(n)
