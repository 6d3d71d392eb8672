"""Batched transport: many independent scenes in one particle pool on one GPU.

The reference runs one scene per ``run_transport`` call (chatcarlo/transport.py:
358-474; its sweeps loop over scenes in Python).  Here every scene of a sweep is
uploaded once and all of their histories go through one pool: a device cursor
over the concatenated history index space refills the pool slots that escaped
or absorbed photons free, and two transport rounds are captured in one hipGraph
that is replayed until the queue is drained.  Scene s's history h draws from the
Philox stream keyed by (seed_s, h), so each scene's kerma and H*(10) grids and
per-material deposits are those of ``run_transport`` with that seed, up to the
order of FP64 additions.

Per scene the batched run keeps totals only (kerma, H*(10), deposits by
material); batch moments (uncertainties) stay with ``run_transport``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

from . import _native as N
from .geometry import Geometry
from .source import SourceRecord
from .tally import VoxelGrid
from .transport import Engine, TransportResult, default_engine


@dataclass
class SceneSpec:
    """One scene of a batch: the scene dict (source, geometry, physics), its dose grid
    (origin_cm (3,), shape (3,), voxel_cm), its history count and its seed."""
    raw: dict
    grid: tuple
    n_histories: int
    seed: int


@dataclass
class Tuning:
    """Launch shape of a batched run: threads per block of the transport kernels, rounds
    between host checks of the queue (even), photons in flight, copies of the scene tallies
    (workgroup b adds into copy b mod copies), and the form: persistent = True runs one
    persistent history-based launch (rounds_per_sync and pool unused) whose workgroups flush
    their LDS tally cache every `epoch` wave iterations."""
    block: int = 256
    rounds_per_sync: int = 8
    pool: int = 1 << 20
    tally_copies: int = 8
    persistent: bool = False
    epoch: int = 32

    def struct(self):
        return N.CcBatchTuning(int(self.block), int(self.rounds_per_sync), int(self.pool), int(self.tally_copies),
                               int(bool(self.persistent)), int(self.epoch))


# the shape of a batched run without a tuning: the persistent form, eight waves per workgroup
# sharing one LDS tally cache, flushed every second iteration, 16 tally copies
PERSISTENT = Tuning(512, tally_copies=16, persistent=True, epoch=2)


@dataclass
class BatchStats:
    n_histories: int
    n_absorbed: int
    n_escaped: int
    scatter_sum: int
    n_fluorescence: int
    n_events: int
    n_flights: int
    n_voxel_visits: int
    rounds: int
    graph_launches: int
    device_ms: float
    tuning: Tuning = field(default_factory=Tuning)

    @property
    def histories_per_s(self):
        return self.n_histories / (self.device_ms * 1e-3) if self.device_ms > 0 else 0.0


class BatchedRun:
    """The scenes of a batch, uploaded to one device context."""

    def __init__(self, scenes: list[SceneSpec], engine: Engine | None = None):
        if not scenes:
            raise ValueError("no scenes")
        self.eng = engine or default_engine()
        self.scenes = list(scenes)
        fl = {bool(s.raw.get("physics", {}).get("fluorescence", True)) for s in self.scenes}
        if len(fl) != 1:
            raise ValueError("all scenes of a batch must agree on physics.fluorescence")
        self.flags = N.RUN_FLUORESCENCE if fl.pop() else 0
        self._upload()
        self.last: BatchStats | None = None
        self._frac = 1.0

    def _upload(self):
        eng = self.eng
        n = len(self.scenes)
        prims, offs, worlds = [], [0], (N.CcWorld * n)()
        self._srcs = [SourceRecord(s.raw["source"]) for s in self.scenes]
        srcs = (N.CcSource * n)()
        for k, s in enumerate(self.scenes):
            if s.raw["source"].get("ctdi_vol_mGy") is not None or s.raw["source"].get("mas") is not None:
                raise NotImplementedError("calibrated sources are not part of the batched path")
            p, w = Geometry(s.raw["geometry"]).packed(eng.pack.mat_id)
            prims.append(p)
            offs.append(offs[-1] + len(p))
            C.memmove(C.byref(worlds[k]), w.ctypes.data, w.nbytes)
            srcs[k] = self._srcs[k].struct
        allp = np.concatenate(prims) if offs[-1] else np.zeros(0, dtype=prims[0].dtype)
        cp = (N.CcPrim * max(1, len(allp)))()
        if len(allp):
            C.memmove(cp, allp.ctypes.data, allp.nbytes)
        off = np.asarray(offs, dtype=np.int32)
        nh = np.asarray([int(s.n_histories) for s in self.scenes], dtype=np.uint64)
        seeds = np.asarray([int(s.seed) & 0xFFFFFFFFFFFFFFFF for s in self.scenes], dtype=np.uint64)
        org = np.ascontiguousarray([np.asarray(s.grid[0], dtype=np.float64) for s in self.scenes])
        shp = np.ascontiguousarray([np.asarray(s.grid[1], dtype=np.int32) for s in self.scenes])
        vox = np.asarray([float(s.grid[2]) for s in self.scenes], dtype=np.float64)
        N.check(eng.lib.cc_batch_upload(eng.ctx, n, cp, N.ptr(off, C.c_int32), worlds, srcs,
                                        N.ptr(nh, C.c_uint64), N.ptr(seeds, C.c_uint64),
                                        N.ptr(org, C.c_double), N.ptr(shp, C.c_int32), N.ptr(vox, C.c_double)))
        eng._scene_key = None   # the single-scene state is untouched, but keep the caches honest
        self._shapes = [tuple(int(x) for x in s.grid[1]) for s in self.scenes]

    def run(self, tuning: Tuning | None = None, fraction: float = 1.0) -> BatchStats:
        """Transports ceil(fraction n_s) histories of every scene; tallies start from zero.  Without
        a tuning: the persistent form (PERSISTENT)."""
        t = tuning or PERSISTENT
        self._frac = float(fraction)
        out = N.CcBatchOut()
        N.check(self.eng.lib.cc_batch_run(self.eng.ctx, C.byref(t.struct()), float(fraction), self.flags,
                                          C.byref(out)))
        self.last = BatchStats(int(out.n_histories), int(out.n_absorbed), int(out.n_escaped),
                               int(out.scatter_sum), int(out.n_fluorescence), int(out.n_events),
                               int(out.n_flights), int(out.n_voxel_visits), int(out.rounds),
                               int(out.graph_launches), float(out.device_ms), t)
        return self.last

    # launch shapes found by autotune, per workload signature (scenes, histories, seeds, grids,
    # fraction, flags): a sweep that comes back is not timed again
    _tuned: dict = {}

    def signature(self, fraction: float) -> str:
        import hashlib
        import json
        h = hashlib.sha256()
        for s in self.scenes:
            h.update(json.dumps(s.raw, sort_keys=True, default=float).encode())
            h.update(repr((int(s.n_histories), int(s.seed), [float(x) for x in np.ravel(s.grid[0])],
                           [int(x) for x in s.grid[1]], float(s.grid[2]))).encode())
        h.update(repr((float(fraction), int(self.flags))).encode())
        return h.hexdigest()

    def autotune(self, fraction: float = 0.02, cached: bool = True):
        """Times launch shapes on a fraction of the histories (three runs over the persistent and
        event-based forms, cc_batch_autotune); returns the fastest and the
        table (persistent, block, epoch, rounds_per_sync, pool, tally_copies, histories/s) of the
        runs.  With cached, a workload tuned before in this process returns its shape and table
        without timing."""
        key = self.signature(fraction)
        if cached and key in BatchedRun._tuned:
            t, table = BatchedRun._tuned[key]
            return replace(t), table.copy()
        best = N.CcBatchTuning()
        table = np.zeros((64, 7))
        rows = C.c_int32()
        N.check(self.eng.lib.cc_batch_autotune(self.eng.ctx, float(fraction), self.flags, C.byref(best),
                                               N.ptr(table, C.c_double), 64, C.byref(rows)))
        tune = Tuning(best.block, best.rounds_per_sync, best.pool, best.tally_copies, bool(best.persistent),
                      int(best.epoch) or 32)
        table = table[:min(rows.value, 64)].copy()
        BatchedRun._tuned[key] = (tune, table)
        return tune, table

    def histories(self, k: int) -> int:
        """Histories of scene k in the last run: ceil(fraction n_k), at least 1."""
        return max(1, int(np.ceil(int(self.scenes[k].n_histories) * self._frac)))

    def read(self, k: int):
        """Scene k's kerma and H*(10) track sums per voxel (the units of VoxelGrid.kerma_keV
        and h10_track_pSv_cm3) and its deposits by material name (keV)."""
        shape = self._shapes[k]
        nv = int(np.prod(shape))
        kk = np.zeros(nv)
        hh = np.zeros(nv)
        ed = np.zeros(N.CC_MAX_MAT)
        N.check(self.eng.lib.cc_batch_read(self.eng.ctx, int(k), N.ptr(kk, C.c_double), N.ptr(hh, C.c_double),
                                           N.ptr(ed, C.c_double)))
        names = self.eng.pack.materials
        edep = {names[m]: float(ed[m]) for m in range(len(names)) if ed[m] != 0.0}
        return kk.reshape(shape), hh.reshape(shape), edep

    def results(self) -> list[TransportResult]:
        """Per-scene results in the reference's result type (grid totals, no moments)."""
        if self.last is None:
            raise RuntimeError("run() first")
        out = []
        for k, s in enumerate(self.scenes):
            kk, hh, edep = self.read(k)
            origin, shape, vox = s.grid
            g = VoxelGrid(origin_cm=np.asarray(origin, dtype=float), shape=tuple(int(x) for x in shape),
                          voxel_size_cm=float(vox), track_uncertainty=False)
            g.kerma_keV[...] = kk
            g.h10_track_pSv_cm3[...] = hh
            n = self.histories(k)
            g.n_histories = n
            out.append(TransportResult(
                n_histories=n, energy_deposited_MeV={m: v / 1000.0 for m, v in edep.items()},
                fraction_absorbed=float("nan"), fraction_escaped=float("nan"), mean_scatter_events=float("nan"),
                grid=g, n_batches=0))
        return out


def run_transport_batch(scenes: list[SceneSpec], tuning: Tuning | None = None, autotune: bool = False,
                        engine: Engine | None = None):
    """Runs every scene of `scenes` in one batched pass; returns (per-scene results, stats)."""
    b = BatchedRun(scenes, engine)
    if autotune:
        tuning, _ = b.autotune()
    st = b.run(tuning)
    return b.results(), st
