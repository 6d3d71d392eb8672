"""Learned importance: an MLP importance map trained on the GPU from analog
contribution records, driving a biased source and weight-window splitting /
Russian roulette, with the analog run as fallback and judge.

The reference's transport is analog (chatcarlo/transport.py:98-233: unit
weights, no splitting); its scoring of H*(10) at a location is the
track-length estimator of chatcarlo/tally.py:248-293.  Here the location is a
sphere (the operator position) scored with the same estimator,
h*(10)(E) x chord x weight / volume, per history.

Flow (all device work is in libchatcarlo_hip.so):

1. ``generate_records``: analog runs with CC_RUN_RECORD.  A sampled subset of
   photons is stored at flight starts (state about to fly: map 0) and at
   collisions (collision point with the incoming direction and energy: map 1),
   each with its score so far; at the batch end the record's target becomes
   the score the photon still collected.  Per flight the expected chord of the
   uncollided flight replaces the realised one in that score (same mean, far
   less noise).
2. ``train``: for each map, Adam on the Poisson deviance exp(f) - s y f,
   f = MLP(state), with the forward/backward GEMMs on BF16 MFMA and FP32
   accumulation; exp(f) estimates the expected future contribution in units of
   s (the same s for both maps).
3. ``configure``: source cells (8^3 cells of the uniforms of field position
   and energy) are drawn with probability proportional to their mean map-0
   importance, mixed with the analog share; weights (1/K)/p_k.  Window centres
   are cw / exp(f), cw = mean source importance.
4. ``run``: transport with CC_RUN_IMPORTANCE | CC_RUN_SOURCE_BIAS until the
   region's relative error reaches the target.  At a collision (map 1) a photon
   is split before its outcome is drawn, so the copies sample independent
   interaction types, energies and directions; at a flight start (map 0) the
   outgoing photon is split or rouletted by where it heads.  Splitting gives n
   copies of w/n; roulette keeps a photon with probability w/wt at weight wt;
   both keep the expected weight, so the estimator stays unbiased.
5. ``compare``: learned and analog runs to the same relative error; figure
   of merit 1/(R^2 t); the learned result is kept only if its mean agrees
   with the analog one within 3 sigma, else the analog result is returned.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .geometry import Geometry
from .source import SourceRecord


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal hidden weights, small output weights, zero biases (float32, CC_MLP_PARAMS)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N.MLP_PARAMS, dtype=np.float32)
    W1 = np.zeros((N.MLP_IN, N.MLP_HIDDEN))
    W1[:7] = rng.normal(0.0, math.sqrt(2.0 / 8.0), (7, N.MLP_HIDDEN))
    o2 = N.MLP_IN * N.MLP_HIDDEN
    w[:o2] = W1.ravel()
    w[o2:o2 + N.MLP_HIDDEN ** 2] = rng.normal(0.0, math.sqrt(2.0 / N.MLP_HIDDEN), N.MLP_HIDDEN ** 2)
    o3 = o2 + N.MLP_HIDDEN ** 2 + N.MLP_HIDDEN
    w[o3:o3 + N.MLP_HIDDEN] = rng.normal(0.0, 0.1 / math.sqrt(N.MLP_HIDDEN), N.MLP_HIDDEN)
    return w


@dataclass
class RegionResult:
    """Region tally of one run: pSv per history and its statistics."""
    n_histories: int
    mean: float
    sem: float
    rel_err: float
    seconds: float
    batch_rel_err: float = float("nan")
    n_split: int = 0
    n_roulette: int = 0
    learned: bool = False
    fallback: bool = False
    kernel_ms: tuple = ()   # source, flight, photo, compton, rayleigh, end_batch, importance (with timing)

    @property
    def fom(self) -> float:
        """Figure of merit 1 / (R^2 t)."""
        if not (self.rel_err > 0) or not (self.seconds > 0):
            return float("nan")
        return 1.0 / (self.rel_err ** 2 * self.seconds)


@dataclass
class TrainResult:
    loss_first: float
    loss_last: float
    ms: float
    yscale: float
    n_records: int
    n_positive: int
    curve: np.ndarray = field(repr=False, default=None)


class _Moments:
    def __init__(self):
        self.n = 0
        self.s = self.s2 = self.b2 = 0.0
        self.nb = 0
        self.split = self.roul = 0
        self.kms = np.zeros(8)

    def add(self, out: N.CcRunOut, n: int):
        self.n += n
        self.s += out.region_sum
        self.s2 += out.region_sum2
        self.b2 += out.region_batch_sum2
        self.nb += out.n_batches
        self.split += out.n_split
        self.roul += out.n_roulette
        self.kms += np.array(out.kernel_ms[:])

    def stats(self):
        n = self.n
        if n < 2:
            return float("nan"), float("nan"), float("nan"), float("nan")
        mean = self.s / n
        var = max(self.s2 / n - mean * mean, 0.0) * n / (n - 1)
        sem = math.sqrt(var / n)
        rel = sem / mean if mean > 0 else float("nan")
        brel = float("nan")
        if self.nb >= 2 and mean > 0:
            bvar = max((self.b2 - self.s * self.s / n) / (self.nb - 1), 0.0)
            brel = math.sqrt(bvar / n) / mean
        return mean, sem, rel, brel


class LearnedImportance:
    """Learned-importance runs for one scene and one spherical scoring region."""

    def __init__(self, engine, scene_raw: dict, center, radius_cm: float, fluorescence: bool = True,
                 feature_scale_cm: float = 100.0):
        self.eng = engine
        self.lib = engine.lib
        self.geo = Geometry(scene_raw["geometry"])
        engine.set_geometry(self.geo)
        self.src = SourceRecord(scene_raw["source"])
        self.center = np.ascontiguousarray(center, dtype=np.float64)
        self.radius = float(radius_cm)
        self.scale = float(feature_scale_cm)
        self.flags = N.RUN_REGION | (N.RUN_FLUORESCENCE if fluorescence else 0)
        self.cw = None
        self.cell_prob = None
        self._trained = False

    def _bind(self):
        self.eng.set_geometry(self.geo)
        N.check(self.lib.cc_set_region(self.eng.ctx, N.ptr(self.center, C.c_double), self.radius, self.scale))

    # -- records -----------------------------------------------------------------
    def generate_records(self, n_histories: int, capacity: int, keep: float = 1.0, seed: int = 1,
                         batch: int = 1_000_000):
        """Analog run that fills the record buffer; returns (records held, CcRunOut)."""
        self._bind()
        N.check(self.lib.cc_records_alloc(self.eng.ctx, int(capacity), float(keep)))
        out = self.eng.run(self.src, int(n_histories), int(seed), 0, int(batch), self.flags | N.RUN_RECORD)
        return int(out.n_records), out

    def read_records(self, kind: int = 0, first: int = 0, n: int | None = None) -> np.ndarray:
        held = C.c_uint32()
        N.check(self.lib.cc_records_read(self.eng.ctx, int(kind), 0, 0, N.fp(), C.byref(held)))
        n = held.value - first if n is None else n
        buf = np.zeros((max(n, 0), N.RECORD_FLOATS), dtype=np.float32)
        if n > 0:
            N.check(self.lib.cc_records_read(self.eng.ctx, int(kind), int(first), int(n), N.ptr(buf, C.c_float),
                                             C.byref(held)))
        return buf

    # -- networks ------------------------------------------------------------------
    def set_params(self, w, map: int = 0):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != N.MLP_PARAMS:
            raise ValueError(f"expected {N.MLP_PARAMS} parameters, got {w.size}")
        N.check(self.lib.cc_mlp_set_params(self.eng.ctx, int(map), N.ptr(w, C.c_float)))

    def get_params(self, map: int = 0) -> np.ndarray:
        w = np.zeros(N.MLP_PARAMS, dtype=np.float32)
        N.check(self.lib.cc_mlp_get_params(self.eng.ctx, int(map), N.ptr(w, C.c_float)))
        return w

    def train(self, map: int = 0, steps: int = 3000, batch: int = 65536, lr: float = 8e-3, lr_end: float = 1e-4,
              seed: int = 0, init_seed: int = 0, yscale: float = 0.0, power: int = 1,
              reinit: bool = True) -> TrainResult:
        if reinit:
            self.set_params(init_params(init_seed + 17 * map), map)
        cfg = N.CcTrainCfg(steps=int(steps), batch=int(batch), lr=float(lr), beta1=0.9, beta2=0.999,
                           eps=1e-8, yscale=float(yscale), lr_end=float(lr_end), seed=int(seed),
                           target_power=int(power))
        out = N.CcTrainOut()
        curve = np.zeros(int(steps))
        N.check(self.lib.cc_mlp_train(self.eng.ctx, int(map), C.byref(cfg), C.byref(out),
                                      N.ptr(curve, C.c_double)))
        self._trained = True
        return TrainResult(out.loss_first, out.loss_last, out.ms, out.yscale, out.n_records, out.n_positive,
                           curve)

    def features(self, pos, dirv, energy) -> np.ndarray:
        self._bind()
        p = N.f32(pos).reshape(-1, 3)
        d = N.f32(dirv).reshape(-1, 3)
        e = N.f32(energy).ravel()
        out = np.zeros((len(e), N.MLP_IN), dtype=np.float32)
        N.check(self.lib.cc_mlp_features(self.eng.ctx, len(e), N.ptr(p, C.c_float), N.ptr(d, C.c_float),
                                         N.ptr(e, C.c_float), N.ptr(out, C.c_float)))
        return out

    def forward(self, feat, map: int = 0) -> np.ndarray:
        x = N.f32(feat).reshape(-1, N.MLP_IN)
        out = np.zeros(len(x), dtype=np.float32)
        N.check(self.lib.cc_mlp_forward(self.eng.ctx, int(map), len(x), N.ptr(x, C.c_float),
                                        N.ptr(out, C.c_float)))
        return out

    # -- importance runs ----------------------------------------------------------------
    def configure(self, w_lo: float = 1e-3, w_hi: float = 3.0, split_max: int = 4, pool_factor: float = 32.0,
                  source_bias: bool = True, bias_mix: float = 0.1, samples_per_cell: int = 64,
                  cw: float = 0.0, use_post: bool = True, use_pre: bool = True, split_gain: float = 1.0,
                  seed: int = 7):
        """Weight windows from the trained map.  w_lo, w_hi bound the window centre (weights
        relative to a unit source photon; a low w_hi keeps an underestimated importance from
        growing heavy photons); split_max caps the copies per split; use_pre splits at
        collisions before their outcome is drawn (k_scatter_split), use_post windows the source
        photons (and every flight start when use_pre is off)."""
        self._bind()
        cfg = N.CcImportanceCfg(cw=float(cw), w_lo=float(w_lo), w_hi=float(w_hi), split_max=int(split_max),
                                pool_factor=float(pool_factor), source_bias=1 if source_bias else 0,
                                bias_mix=float(bias_mix), samples_per_cell=int(samples_per_cell),
                                use_post=1 if use_post else 0, use_pre=1 if use_pre else 0,
                                split_gain=float(split_gain), seed=int(seed))
        cw_out = C.c_float()
        prob = np.zeros(N.SRC_CELLS, dtype=np.float32)
        N.check(self.lib.cc_set_importance(self.eng.ctx, C.byref(cfg), C.byref(self.src.struct),
                                           C.byref(cw_out), N.ptr(prob, C.c_float)))
        self.cw = cw_out.value
        self.cell_prob = prob if source_bias else None
        self._source_bias = bool(source_bias)
        return self.cw

    def run(self, learned: bool, target_rel_err: float | None = 0.02, n_histories: int | None = None,
            seed: int = 11, batch: int = 1_000_000, chunk_batches: int = 1, min_histories: int = 500_000,
            max_histories: int = 2_000_000_000, max_seconds: float = 120.0, timing: bool = False) -> RegionResult:
        """Run until the region's relative error reaches target_rel_err (or exactly n_histories)."""
        self._bind()
        flags = self.flags | (N.RUN_TIMING if timing else 0)
        if learned:
            if self.cw is None:
                raise RuntimeError("configure() the importance map before a learned run")
            flags |= N.RUN_IMPORTANCE | (N.RUN_SOURCE_BIAS if self._source_bias else 0)
        mom = _Moments()
        chunk = batch * chunk_batches
        first = 0
        t0 = time.perf_counter()
        while True:
            n = chunk if n_histories is None else min(chunk, n_histories - first)
            if n <= 0:
                break
            out = self.eng.run(self.src, n, seed, first, batch, flags)
            mom.add(out, n)
            first += n
            if n_histories is None:
                _, _, rel, _ = mom.stats()
                if first >= min_histories and rel == rel and rel <= target_rel_err:
                    break
                if first >= max_histories or time.perf_counter() - t0 > max_seconds:
                    break
        dt = time.perf_counter() - t0
        mean, sem, rel, brel = mom.stats()
        return RegionResult(mom.n, mean, sem, rel, dt, brel, mom.split, mom.roul, learned=learned,
                            kernel_ms=tuple(round(float(x), 2) for x in mom.kms[:7]) if timing else ())

    def compare(self, target_rel_err: float = 0.02, seed_analog: int = 101, seed_learned: int = 202,
                batch: int = 1_000_000, **kw):
        """Analog and learned runs to the same relative error.  Returns (analog, learned, chosen, z):
        z is the difference of the means in combined standard errors.  The analog run judges the
        learned one: its result is chosen only if |z| <= 3 and its figure of merit is higher."""
        a = self.run(False, target_rel_err, seed=seed_analog, batch=batch, **kw)
        b = self.run(True, target_rel_err, seed=seed_learned, batch=batch, **kw)
        z = (b.mean - a.mean) / math.sqrt(a.sem ** 2 + b.sem ** 2)
        chosen = b if abs(z) <= 3.0 and b.fom > a.fom else a
        if chosen is a:
            a.fallback = True
        return a, b, chosen, z
