"""Oracle of the learned importance map: the 7 -> 64 -> 64 -> 1 MLP, its Poisson
loss and gradient, Adam, the feature map, and the weight-window rules.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference's transport is analog only (chatcarlo/transport.py:98-233 has no
weights, splitting or roulette), so nothing here restates a reference
function; these are the published algorithms the GPU path implements:

* f(x) = w3 . relu(relu(x W1) W2 + b2) + b3 on 16 padded inputs (7 features,
  a constant 1 whose W1 row is the first-layer bias, zeros);
* Poisson deviance  L = sum_r g (exp(f_r) - s y_r f_r)  (the negative
  log-likelihood of y ~ Poisson(exp f) up to terms free of f), whose minimiser
  is f = ln E[s y | x];
* Adam (Kingma and Ba, ICLR 2015, algorithm 1);
* weight windows with splitting and Russian roulette (MCNP-style, Booth 1985).

``bf16=True`` rounds the operands the GPU stores in BF16 (inputs, W1, W2,
relu(Z1), and the backward deltas D2, D1) to nearest-even BF16 and keeps every
other quantity in FP64; ``bf16=False`` is the plain FP64 map.

Pinning: the FP64 gradient is checked against central finite differences of
the FP64 loss (tests/test_oracle_mlp.py), which a dropped term, a sign error
or a transposed weight fails; forward is checked against a per-sample loop
over the scalar definition; Adam against its closed form for a constant
gradient; the window rules by exact expectations (sum of split weights equals
the parent's, roulette keeps E[w]); the feature map by hand-computed rows.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 16, 64
OFF_W1 = 0
OFF_W2 = OFF_W1 + N_IN * N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H
N_PARAMS = OFF_B3 + 1


def bf16_round(x):
    """Nearest-even BF16 of float32(x), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    W1 = w[OFF_W1:OFF_W2].reshape(N_IN, N_H).copy()
    W1[8:] = 0.0
    W2 = w[OFF_W2:OFF_B2].reshape(N_H, N_H)
    return W1, W2, w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3], float(w[OFF_B3])


def init_params(seed=0):
    """He-normal W1 (feature and bias rows), W2; w3 small; zero biases."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N_PARAMS)
    W1 = np.zeros((N_IN, N_H))
    W1[:8] = rng.normal(0.0, np.sqrt(2.0 / 8.0), (8, N_H))
    W1[7] = 0.0
    w[OFF_W1:OFF_W2] = W1.ravel()
    w[OFF_W2:OFF_B2] = rng.normal(0.0, np.sqrt(2.0 / N_H), N_H * N_H)
    w[OFF_W3:OFF_B3] = rng.normal(0.0, 0.1 / np.sqrt(N_H), N_H)
    return w.astype(np.float32)


def features(pos, dirv, energy, centre, scale):
    """Feature rows in the frame of the scoring region: (p - c) / scale; the direction in the
    basis u = unit(c - p), e2 = unit(z x u) (unit(u x x) when |u_z| >= 0.999), e3 = u x e2;
    ln(E / 50 keV) / 2; 1; zeros."""
    c = np.asarray(centre, dtype=np.float64)
    pos = np.atleast_2d(np.asarray(pos, dtype=np.float64))
    d = np.atleast_2d(np.asarray(dirv, dtype=np.float64))
    n = pos.shape[0]
    x = np.zeros((n, N_IN))
    r = c - pos
    x[:, 0:3] = -r / scale
    rl = np.linalg.norm(r, axis=1)
    u = np.where(rl[:, None] > 0, r / np.where(rl > 0, rl, 1.0)[:, None], np.array([0.0, 0.0, 1.0]))
    polar = np.abs(u[:, 2]) >= 0.999
    e2 = np.stack([-u[:, 1], u[:, 0], np.zeros(n)], axis=1)
    e2[polar] = np.stack([np.zeros(polar.sum()), u[polar, 2], -u[polar, 1]], axis=1)
    e2 /= np.linalg.norm(e2, axis=1, keepdims=True)
    e3 = np.cross(u, e2)
    x[:, 3] = np.einsum("ij,ij->i", d, u)
    x[:, 4] = np.einsum("ij,ij->i", d, e2)
    x[:, 5] = np.einsum("ij,ij->i", d, e3)
    x[:, 6] = 0.5 * np.log(np.asarray(energy, dtype=np.float64) / 50.0)
    x[:, 7] = 1.0
    return x


def forward(w, x, bf16=True):
    """f per row, with the intermediates the backward pass needs."""
    W1, W2, b2, w3, b3 = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    if bf16:
        x, W1, W2 = bf16_round(x), bf16_round(W1), bf16_round(W2)
    z1 = x @ W1
    h1 = np.maximum(z1, 0.0)
    if bf16:
        h1 = bf16_round(h1)
    z2 = h1 @ W2 + b2
    f = np.maximum(z2, 0.0) @ w3 + b3
    return f, dict(x=x, W1=W1, W2=W2, z1=z1, h1=h1, z2=z2, w3=w3)


def forward_loop(w, x):
    """Scalar per-sample restatement of the FP64 forward (for pinning `forward`)."""
    W1, W2, b2, w3, b3 = unpack(w)
    out = []
    for row in np.asarray(x, dtype=np.float64):
        h1 = [max(sum(row[a] * W1[a, i] for a in range(N_IN)), 0.0) for i in range(N_H)]
        s = b3
        for j in range(N_H):
            z = b2[j] + sum(h1[i] * W2[i, j] for i in range(N_H))
            s += max(z, 0.0) * w3[j]
        out.append(s)
    return np.array(out)


def loss(w, x, y, yscale=1.0, gscale=1.0, bf16=False):
    f, _ = forward(w, x, bf16)
    ys = yscale * np.asarray(y, dtype=np.float64)
    return float(np.sum((np.exp(np.minimum(f, 80.0)) - ys * f) * gscale))


def grad(w, x, y, yscale=1.0, gscale=1.0, bf16=True):
    """Gradient of `loss` with respect to the flat parameters (rows 8..15 of W1 zero)."""
    f, t = forward(w, x, bf16)
    ys = yscale * np.asarray(y, dtype=np.float64)
    g = (np.exp(np.minimum(f, 80.0)) - ys) * gscale
    z2, z1, h1, w3 = t["z2"], t["z1"], t["h1"], t["w3"]
    out = np.zeros(N_PARAMS)
    out[OFF_W3:OFF_B3] = np.maximum(z2, 0.0).T @ g
    out[OFF_B3] = g.sum()
    d2 = (g[:, None] * w3[None, :]) * (z2 > 0)
    out[OFF_B2:OFF_W3] = d2.sum(axis=0)
    if bf16:
        d2 = bf16_round(d2)
    out[OFF_W2:OFF_B2] = (h1.T @ d2).ravel()
    d1 = (d2 @ t["W2"].T) * (z1 > 0)
    if bf16:
        d1 = bf16_round(d1)
    gw1 = t["x"].T @ d1
    gw1[8:] = 0.0
    out[OFF_W1:OFF_W2] = gw1.ravel()
    return out


def adam_step(w, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); rows 8..15 of W1 stay zero."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    w = w - lr * (m / (1.0 - b1 ** t)) / (np.sqrt(v / (1.0 - b2 ** t)) + eps)
    W1 = w[OFF_W1:OFF_W2].reshape(N_IN, N_H)
    W1[8:] = 0.0
    return w, m, v


def target_weight(f, cw, w_lo, w_hi):
    """Weight-window centre clamp(cw / exp(f), w_lo, w_hi)."""
    f = np.clip(np.asarray(f, dtype=np.float64), -80.0, 80.0)
    return np.clip(cw * np.exp(-f), w_lo, w_hi)


def window(w, wt, split_max, u):
    """Weight window [wt / 2, 2 wt]: returns (copies n, weight per copy, survived).
    Above: n = min(floor(w / wt), split_max) copies of w / n.  Below: survive
    with probability w / wt (uniform u) at weight wt.  Inside: unchanged."""
    w = np.asarray(w, dtype=np.float64)
    wt = np.asarray(wt, dtype=np.float64)
    n = np.ones_like(w, dtype=np.int64)
    out_w = w.copy()
    alive = np.ones_like(w, dtype=bool)
    hi = w > 2.0 * wt
    n[hi] = np.minimum(np.floor(w[hi] / wt[hi]).astype(np.int64), split_max)
    out_w[hi] = w[hi] / n[hi]
    lo = w < 0.5 * wt
    alive[lo] = np.asarray(u)[lo] < w[lo] / wt[lo]
    out_w[lo] = np.where(alive[lo], wt[lo], 0.0)
    return n, out_w, alive


def source_cells(imp, mix):
    """Cell probabilities p_k = (1 - mix) imp_k / sum imp + mix / K and weights (1 / K) / p_k."""
    imp = np.asarray(imp, dtype=np.float64)
    K = imp.size
    p = (1.0 - mix) * imp / imp.sum() + mix / K
    return p, (1.0 / K) / p
