"""GPU learned importance: MFMA MLP kernels against the oracle, the region tally
against the oracle transport, contribution records, exact splitting / roulette,
source biasing, and the trained map's unbiasedness and figure-of-merit gain
over the analog run."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import core
from oracle import mlp as M
from oracle import transport as OT

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2027)


def _lib(engine):
    return engine.lib, engine.ctx


def _random_params(seed=3):
    rng = np.random.default_rng(seed)
    w = M.init_params(seed).astype(np.float64)
    W1 = w[M.OFF_W1:M.OFF_W2].reshape(M.N_IN, M.N_H)
    W1[7] = rng.normal(0, 0.3, M.N_H)
    w[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.2, M.N_H)
    w[M.OFF_W3:M.OFF_B3] = rng.normal(0, 0.3, M.N_H)
    w[M.OFF_B3] = -0.4
    return w.astype(np.float32)


def _rows(n, rng):
    x = np.zeros((n, M.N_IN), dtype=np.float32)
    x[:, :3] = rng.uniform(-1, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    x[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    x[:, 6] = 0.5 * np.log(rng.uniform(2, 150, n) / 50.0)
    x[:, 7] = 1.0
    return x


def _set_params(engine, w):
    from chatcarlo_amd import _native as N
    lib, ctx = _lib(engine)
    w = np.ascontiguousarray(w, dtype=np.float32)
    N.check(lib.cc_mlp_set_params(ctx, 0, N.ptr(w, C.c_float)))


def _forward_scale(w, x):
    """Sum of |terms| of the output layer per row: the scale of f's rounding error."""
    f, t = M.forward(w.astype(np.float64), x, bf16=True)
    return np.abs(np.maximum(t["z2"], 0.0) * t["w3"]).sum(axis=1) + abs(float(w[M.OFF_B3]))


@pytest.mark.parametrize("n", [1, 64, 1000])
def test_mfma_forward_matches_oracle(engine, n):
    from chatcarlo_amd import _native as N
    lib, ctx = _lib(engine)
    w = _random_params(11)
    _set_params(engine, w)
    x = _rows(n, RNG)
    out = np.zeros(n, dtype=np.float32)
    N.check(lib.cc_mlp_forward(ctx, 0, n, N.ptr(x, C.c_float), N.ptr(out, C.c_float)))
    ref, _ = M.forward(w.astype(np.float64), x, bf16=True)
    scale = _forward_scale(w, x)
    err = np.abs(out - ref) / (1.0 + scale)
    # FP32 accumulation against FP64: 1e-5 of the term scale; a BF16 rounding of relu(Z1)
    # that falls on the other side of a tie moves a row by up to ~4e-3 of one term
    assert np.quantile(err, 0.99) < 1e-5 if n >= 100 else err.max() < 1e-4
    assert err.max() < 2e-3
    # the unrounded FP64 map differs at the BF16 level only
    ref64, _ = M.forward(w.astype(np.float64), x, bf16=False)
    assert np.max(np.abs(out - ref64) / (1.0 + scale)) < 0.05


@pytest.mark.parametrize("n,yscale", [(32, 1.0), (1000, 1.7)])
def test_mlp_gradient_matches_oracle(engine, n, yscale):
    from chatcarlo_amd import _native as N
    lib, ctx = _lib(engine)
    w = _random_params(5)
    _set_params(engine, w)
    x = _rows(n, RNG)
    y = RNG.poisson(0.8, n).astype(np.float32)
    g = np.zeros(N.MLP_PARAMS, dtype=np.float32)
    loss = C.c_double()
    N.check(lib.cc_mlp_grad(ctx, 0, n, N.ptr(x, C.c_float), N.ptr(y, C.c_float), float(yscale),
                            N.ptr(g, C.c_float), C.byref(loss)))
    ref = M.grad(w.astype(np.float64), x, y, yscale=yscale, bf16=True)
    ref_loss = M.loss(w.astype(np.float64), x, y, yscale, bf16=True)
    assert abs(loss.value - ref_loss) <= 1e-5 * (abs(ref_loss) + n)
    blocks = {"W1": slice(M.OFF_W1, M.OFF_W2), "W2": slice(M.OFF_W2, M.OFF_B2), "b2": slice(M.OFF_B2, M.OFF_W3),
              "w3": slice(M.OFF_W3, M.OFF_B3), "b3": slice(M.OFF_B3, M.N_PARAMS)}
    for name, sl in blocks.items():
        rel = np.linalg.norm(g[sl] - ref[sl]) / max(np.linalg.norm(ref[sl]), 1e-30)
        assert rel < 3e-3, (name, rel)
    assert np.all(g[M.OFF_W1 + 8 * M.N_H:M.OFF_W2] == 0)


def _angio(engine):
    from chatcarlo_amd.importance import LearnedImportance
    from chatcarlo_amd.workloads import angio_room, angio_region
    raw, _ = angio_room()
    c, r = angio_region()
    return LearnedImportance(engine, raw, c, r)


def test_features_match_oracle(engine):
    li = _angio(engine)
    n = 500
    pos = RNG.uniform(-200, 200, (n, 3)).astype(np.float32)
    d = RNG.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    e = RNG.uniform(1, 150, n).astype(np.float32)
    got = li.features(pos, d, e)
    li._bind()
    ref = M.features(pos, d, e, li.center, li.scale)
    np.testing.assert_allclose(got, ref, atol=2e-6, rtol=2e-6)


def _side_scene():
    """60 keV 2x2 cm parallel beam through a 20 cm water cube; scoring sphere beside the cube."""
    raw = {
        "source": {"position": [0.0, 0.0, -40.0], "direction": [0.0, 0.0, 1.0],
                   "field": {"shape": "parallel", "size_cm": [2.0, 2.0]},
                   "spectrum": [{"energy_keV": 60.0, "weight": 1.0}]},
        "geometry": [
            {"name": "room", "shape": "box", "material": "air", "center": [0, 0, 0], "size_cm": [100, 100, 100]},
            {"name": "cube", "shape": "box", "material": "water", "center": [0, 0, 0], "size_cm": [20, 20, 20]},
        ],
        "physics": {"fluorescence": True},
    }
    return raw, np.array([20.0, 0.0, 0.0]), 8.0


def test_region_tally_matches_oracle(engine, dpack):
    from chatcarlo_amd.importance import LearnedImportance
    raw, c, r = _side_scene()
    li = LearnedImportance(engine, raw, c, r)
    gpu = li.run(False, None, n_histories=4_000_000, seed=3, batch=100_000, chunk_batches=10)
    prims, world = li.geo.packed(dpack.mat_id)
    og = core.OGeom.from_packed(prims, world)
    n_or = 100_000
    res = OT.run(raw["source"], og, dpack, li.src.spec_e, li.src.spec_cdf, n_or, 9, batch_size=5_000,
                 region=(c, r))
    m_or = res["region_sum"] / n_or
    se_or = math.sqrt(max(res["region_sum2"] / n_or - m_or ** 2, 0.0) / (n_or - 1))
    assert gpu.mean > 0 and m_or > 0
    z = (gpu.mean - m_or) / math.sqrt(gpu.sem ** 2 + se_or ** 2)
    assert abs(z) < 4.0, (gpu.mean, m_or, z)
    # history and batch estimates of the error agree (40 batches)
    assert 0.6 < gpu.rel_err / gpu.batch_rel_err < 1.6


def test_records_hold_the_future_score(engine):
    from chatcarlo_amd.importance import LearnedImportance
    raw, c, r = _side_scene()
    li = LearnedImportance(engine, raw, c, r)
    n_rec, out = li.generate_records(200_000, 4_000_000, keep=1.0, seed=5, batch=50_000)
    assert 0 < n_rec < 4_000_000
    rec = li.read_records(kind=0)
    assert len(rec) + len(li.read_records(kind=1)) == n_rec
    slot = rec[:, 9].view(np.int32)
    batch = rec[:, 10].view(np.uint32)
    assert np.all(rec[:, 8] >= 0) and np.all(rec[:, 7] >= 0)
    key = batch.astype(np.int64) * (1 << 32) + slot
    _, first = np.unique(key, return_index=True)
    # one round-0 record per history, with nothing scored before it
    assert len(first) == 200_000
    assert np.all(rec[first, 7] == 0)
    # its target is the whole history's expected score: flight by flight, the expected chord of
    # the uncollided flight replaces the realised one, so the mean is the region tally's and the
    # variance is far smaller
    y0 = rec[first, 8].astype(np.float64)
    n = len(y0)
    var_y = y0.var(ddof=1)
    var_r = (out.region_sum2 - out.region_sum ** 2 / n) / (n - 1)
    z = (y0.sum() - out.region_sum) / math.sqrt(n * (var_y + var_r))
    assert abs(z) < 4.0, (y0.sum(), out.region_sum, z)
    assert var_y < 0.7 * var_r
    # score before + target is the same final score for every record of a history
    final = rec[:, 7].astype(np.float64) + rec[:, 8]
    order = np.argsort(key, kind="stable")
    ks, fs = key[order], final[order]
    starts = np.r_[0, np.nonzero(np.diff(ks))[0] + 1]
    ref = np.repeat(fs[starts], np.diff(np.r_[starts, len(ks)]))
    assert np.all(np.abs(fs - ref) <= 1e-5 * np.maximum(ref, 1e-30) + 1e-12)
    # collision records: same history keys, nothing is scored between a collision and the next
    # flight start, and the final score of the history is the same
    col = li.read_records(kind=1)
    assert len(col) > 0 and np.all(col[:, 11] == 1) and np.all(rec[:, 11] == 0)
    ck = col[:, 10].view(np.uint32).astype(np.int64) * (1 << 32) + col[:, 9].view(np.int32)
    assert np.all(np.isin(ck, key))
    fin = dict(zip(ks[starts].tolist(), fs[starts].tolist()))
    cf = col[:, 7].astype(np.float64) + col[:, 8]
    want = np.array([fin[k] for k in ck.tolist()])
    assert np.all(np.abs(cf - want) <= 1e-5 * np.maximum(want, 1e-30) + 1e-12)


def _constant_map(engine, li, f_const, cw, split_max=16, pool_factor=9.0, post=True, pre=False):
    w = np.zeros(M.N_PARAMS, dtype=np.float32)
    w[M.OFF_B3] = f_const
    li.set_params(w, 0)
    li.set_params(w, 1)
    li.configure(w_lo=1e-3, w_hi=1e3, split_max=split_max, pool_factor=pool_factor, source_bias=False,
                 cw=cw, use_post=post, use_pre=pre)


def test_splitting_is_exact_and_unbiased(engine):
    from chatcarlo_amd.importance import LearnedImportance
    raw, c, r = _side_scene()
    li = LearnedImportance(engine, raw, c, r)
    ana = li.run(False, None, n_histories=2_000_000, seed=21, batch=200_000, chunk_batches=5)
    _constant_map(engine, li, 0.0, 1.0 / 8.0)
    n = 1_000_000
    spl = li.run(True, None, n_histories=n, seed=22, batch=100_000, chunk_batches=5)
    # every source photon (weight 1 > 2 wt) becomes 8 photons of weight 1/8, which stay in the window
    assert spl.n_split == 7 * n
    assert spl.n_roulette == 0
    z = (spl.mean - ana.mean) / math.sqrt(spl.sem ** 2 + ana.sem ** 2)
    assert abs(z) < 4.0, (spl.mean, ana.mean, z)
    assert spl.rel_err < ana.rel_err   # 8 correlated-at-birth copies still beat one photon per history


def test_collision_splitting_is_unbiased(engine):
    """Split at the first collision, before the outcome is drawn: 8 copies of weight 1/8 that
    each draw their own interaction type, energy and direction (and deposit their share)."""
    from chatcarlo_amd.importance import LearnedImportance
    raw, c, r = _side_scene()
    li = LearnedImportance(engine, raw, c, r)
    ana = li.run(False, None, n_histories=2_000_000, seed=23, batch=200_000, chunk_batches=5)
    _constant_map(engine, li, 0.0, 1.0 / 8.0, post=False, pre=True)
    n = 1_000_000
    spl = li.run(True, None, n_histories=n, seed=24, batch=100_000, chunk_batches=5)
    # first collisions split into 8 copies of weight 1/8 (later ones sit at the window centre);
    # a copy whose azimuth can reach the region leaves as two strata, the lighter one below the
    # window, so roulette acts too
    assert spl.n_split > 0 and spl.n_roulette > 0
    z = (spl.mean - ana.mean) / math.sqrt(spl.sem ** 2 + ana.sem ** 2)
    assert abs(z) < 4.0, (spl.mean, ana.mean, z)
    # independent outcomes per copy lower the variance per history
    assert spl.rel_err * math.sqrt(spl.n_histories) < 0.8 * ana.rel_err * math.sqrt(ana.n_histories)


def test_roulette_is_unbiased(engine):
    from chatcarlo_amd.importance import LearnedImportance
    raw, c, r = _side_scene()
    li = LearnedImportance(engine, raw, c, r)
    ana = li.run(False, None, n_histories=2_000_000, seed=31, batch=200_000, chunk_batches=5)
    _constant_map(engine, li, 0.0, 8.0)
    n = 4_000_000
    rr = li.run(True, None, n_histories=n, seed=32, batch=200_000, chunk_batches=5)
    assert rr.n_split == 0
    p = 7.0 / 8.0
    assert abs(rr.n_roulette - p * n) < 5 * math.sqrt(n * p * (1 - p))
    z = (rr.mean - ana.mean) / math.sqrt(rr.sem ** 2 + ana.sem ** 2)
    assert abs(z) < 4.0, (rr.mean, ana.mean, z)


def test_biased_source_alone_is_unbiased(engine):
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.importance import LearnedImportance, _Moments
    raw, c, r = _side_scene()
    li = LearnedImportance(engine, raw, c, r)
    li.set_params(_random_params(8))
    li.configure(source_bias=True, bias_mix=0.2, samples_per_cell=16, use_pre=False)
    p = li.cell_prob.astype(np.float64)
    assert abs(p.sum() - 1) < 1e-5 and p.min() >= 0.2 / 512 * (1 - 1e-5)
    ana = li.run(False, None, n_histories=2_000_000, seed=41, batch=200_000, chunk_batches=5)
    mom = _Moments()
    for k in range(5):
        out = engine.run(li.src, 400_000, 42, 400_000 * k, 200_000,
                         N.RUN_REGION | N.RUN_FLUORESCENCE | N.RUN_SOURCE_BIAS)
        mom.add(out, 400_000)
    mean, sem, _, _ = mom.stats()
    z = (mean - ana.mean) / math.sqrt(sem ** 2 + ana.sem ** 2)
    assert abs(z) < 4.0, (mean, ana.mean, z)


def test_adam_step_matches_oracle(engine):
    from chatcarlo_amd.importance import LearnedImportance
    raw, c, r = _side_scene()
    li = LearnedImportance(engine, raw, c, r)
    n_rec, _ = li.generate_records(20_000, 1_000_000, keep=0.25, seed=51, batch=20_000)
    rec = li.read_records(kind=0).astype(np.float64)
    n_rec = len(rec)
    w0 = _random_params(9)
    li.set_params(w0)
    res = li.train(0, steps=1, batch=n_rec, lr=1e-3, reinit=False)
    w1 = li.get_params(0).astype(np.float64)
    x = M.features(rec[:, 0:3], rec[:, 3:6], rec[:, 6], li.center, li.scale)
    assert res.n_records == n_rec and res.n_positive > 0
    ys = n_rec / rec[:, 8].sum()
    assert abs(res.yscale - ys) <= 1e-5 * ys
    g = M.grad(w0.astype(np.float64), x, rec[:, 8], yscale=ys, gscale=1.0 / n_rec, bf16=True)
    wr, _, _ = M.adam_step(w0.astype(np.float64), np.zeros(M.N_PARAMS), np.zeros(M.N_PARAMS), g, 1, lr=1e-3)
    dw, dr = w1 - w0, wr - w0
    # the first Adam step is lr g / (|g| + eps'): components with a clear gradient move by lr sign(g)
    big = np.abs(g) > 1e-2 * np.abs(g).max()
    np.testing.assert_allclose(dw[big], dr[big], rtol=2e-3, atol=2e-6)
    assert np.all(np.abs(dw) <= 1e-3 * (1 + 1e-4))
    assert np.all(w1[M.OFF_W1 + 8 * M.N_H:M.OFF_W2] == 0)


def test_training_is_reproducible(engine):
    """Each gradient block writes its partial sum and the Adam kernel adds the partials in one
    fixed order, so two trainings from the same parameters on the same records give bit-identical
    weights and losses."""
    from chatcarlo_amd.importance import LearnedImportance
    raw, c, r = _side_scene()
    li = LearnedImportance(engine, raw, c, r)
    n_rec, _ = li.generate_records(50_000, 1_000_000, keep=0.5, seed=52, batch=50_000)
    assert n_rec > 4096
    runs = []
    for _ in range(2):
        li.set_params(_random_params(13))
        res = li.train(0, steps=20, batch=4096, lr=3e-3, reinit=False)
        runs.append((li.get_params(0), res.loss_first, res.loss_last))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][1:] == runs[1][1:]
    assert not np.array_equal(runs[0][0], _random_params(13))


def test_learned_importance_is_unbiased_and_beats_analog(engine):
    """Trained map on the angio room: every learned run agrees with its analog run within 3.5 sigma,
    and the median figure of merit 1/(R^2 t) over nine seed pairs, at the same relative error, is at
    least ten times the analog one (the learned-importance target, training excluded).  Single
    pairs spread by a factor of a few (heavy-tailed weights: 4.8 to 17.6 in the pairs measured so
    far, about one in seven below 10), so the bar is on the median of nine, which falls below 10
    with a probability under 1 % at that spread."""
    li = _angio(engine)
    n_rec, _ = li.generate_records(10_000_000, 32_000_000, keep=0.3, seed=61)
    t0 = li.train(0, steps=3000, batch=65536, lr=8e-3, lr_end=1e-4)
    assert t0.loss_last < t0.loss_first
    li.configure()
    gains = []
    for k in range(9):
        ana, lrn, chosen, z = li.compare(target_rel_err=0.02, seed_analog=71 + 2 * k, seed_learned=72 + 2 * k)
        print(f"pair {k}: analog R={ana.rel_err:.4f} t={ana.seconds:.3f}s FOM={ana.fom:.1f} | learned "
              f"R={lrn.rel_err:.4f} t={lrn.seconds:.3f}s FOM={lrn.fom:.1f} gain={lrn.fom / ana.fom:.2f} z={z:.2f}")
        assert abs(z) < 3.5   # nine pairs: 3.5 sigma keeps a false alarm near 0.4 %
        gains.append(lrn.fom / ana.fom)
    assert sorted(gains)[4] > 10.0, gains
