"""Photon transport on the GPU: run_transport / transport_photons (drop-in API).

Reference: chatcarlo/transport.py:88-474.  The whole history loop runs in
libchatcarlo_hip.so (event-based kernels, see csrc/chatcarlo_hip.hip); this
module resolves the scene into packed records, owns the device context and
turns the C-ABI results back into the reference's result types.

Differences from the reference that a caller can see:
* the RNG is Philox4x32-10 keyed by (seed, history index), so a result depends
  on (seed, first history) only -- not on batch size, n_workers or device
  count; the reference's PCG64 stream is not reproduced (statistical parity);
* n_workers is accepted and ignored (one GPU replaces process-level workers);
* tallies are scored by an exact DDA (the expectation of the reference's
  stratified estimator) instead of stratified random sub-steps.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .datapack import build_pack
from .geometry import Geometry
from .source import SourceRecord
from .tally import ScalarMoments, VoxelGrid


@dataclass
class BatchResult:
    """chatcarlo/transport.py:88-95."""
    n_scatter: np.ndarray
    absorbed: np.ndarray
    escaped: np.ndarray
    final_energy: np.ndarray
    energy_deposited: dict = field(default_factory=dict)
    n_fluorescence: int = 0


@dataclass
class TransportResult:
    """chatcarlo/transport.py:236-254."""
    n_histories: int
    energy_deposited_MeV: dict
    fraction_absorbed: float
    fraction_escaped: float
    mean_scatter_events: float
    grid: VoxelGrid | None = None
    n_photons_real: float | None = None
    n_fluorescence: int = 0
    n_batches: int = 0
    energy_deposited_sem_MeV: dict = field(default_factory=dict)
    energy_deposited_rel_err: dict = field(default_factory=dict)
    device_stats: dict = field(default_factory=dict)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Engine:
    """One device context with the data pack uploaded."""

    def __init__(self, device: int = 0):
        self.lib = N.load()
        ctx = C.c_void_p()
        N.check(self.lib.cc_ctx_create(int(device), C.byref(ctx)))
        self.ctx = ctx
        self.pack = build_pack()
        self._keep = []
        self._upload_tables()
        self._scene_key = None
        self._grid_key = None

    def close(self):
        if self.ctx:
            self.lib.cc_ctx_destroy(self.ctx)
            self.ctx = None

    # -- tables ---------------------------------------------------------------
    def _upload_tables(self):
        p = self.pack
        n_mat, n_e = len(p.materials), len(p.e_node)
        t = N.CcTables()
        keep = {}

        def fa(name, arr):
            a = _f32(arr).ravel()
            keep[name] = a
            return N.ptr(a, C.c_float)

        def ia(name, arr):
            a = np.ascontiguousarray(arr, dtype=np.int32).ravel()
            keep[name] = a
            return N.ptr(a, C.c_int32)

        t.n_mat, t.n_e, t.n_el, t.n_q = n_mat, n_e, len(p.elements), len(p.q_s)
        t.max_edges, t.n_h10 = p.edge_e.shape[1], len(p.h10_e)
        t.density = fa("density", p.density)
        t.e_node = fa("e_node", p.e_node)
        t.dl_node = fa("dl_node", p.dl_node)
        t.sig = fa("sig", p.sig)
        t.slope = fa("slope", p.slope)
        t.edge_of_interval = ia("eoi", p.edge_of_interval)
        t.edge_e = fa("edge_e", p.edge_e)
        t.edge_dl = fa("edge_dl", p.edge_dl)
        t.edge_sig = fa("edge_sig", p.edge_sig)
        t.edge_slope = fa("edge_slope", p.edge_slope)
        off = np.zeros(n_mat + 1, dtype=np.int32)
        for m in range(n_mat):
            off[m + 1] = off[m] + len(p.mat_elem[m])
        t.mat_elem_off = ia("meo", off)
        me = np.concatenate([np.asarray(x, dtype=np.int32) for x in p.mat_elem])
        t.mat_elem = ia("me", me)
        t.elem_len = len(me)
        sel_off = np.zeros(n_mat, dtype=np.int32)
        chunks, acc = [], 0
        for m in range(n_mat):
            sel_off[m] = acc
            c = _f32(p.sel_cdf[m]).ravel()
            chunks.append(c)
            acc += len(c)
        t.sel_off = ia("so", sel_off)
        t.sel = fa("sel", np.concatenate(chunks))
        t.sel_len = acc
        t.sq_over_z = fa("sq", p.sq_over_z)
        t.q_s = fa("qs", p.q_s)
        t.ray_x = fa("rx", p.ray_x)
        t.ray_a = fa("ra", p.ray_a)
        fl = np.zeros((len(p.elements), 32))
        fl[:, 0], fl[:, 1], fl[:, 2] = p.fl_edge, p.fl_omega, p.fl_kinv
        fl[:, 3] = np.log(np.where(p.fl_edge > 0, p.fl_edge, 1.0))
        fl[:, 4:12], fl[:, 12:20], fl[:, 20:32] = p.fl_lines, p.fl_cdf, p.fl_kfrac
        t.fl = fa("fl", fl)
        nh = len(p.h10_e)
        h = np.zeros((4, nh))
        h[0], h[1] = p.h10_e, p.h10_v
        h[2, :nh - 1], h[3, :nh - 1] = p.h10_dl, p.h10_slope
        t.h10 = fa("h10", h)
        self._tables_keep = keep
        N.check(self.lib.cc_upload_tables(self.ctx, C.byref(t)))

    # -- scene / grid -----------------------------------------------------------
    def set_geometry(self, geometry: Geometry):
        prims, world = geometry.packed(self.pack.mat_id)
        key = prims.tobytes() + world.tobytes()
        if key == self._scene_key:
            return prims, world
        cp = (N.CcPrim * max(1, len(prims)))()
        C.memmove(cp, prims.ctypes.data, prims.nbytes)
        w = N.CcWorld()
        C.memmove(C.byref(w), world.ctypes.data, world.nbytes)
        N.check(self.lib.cc_upload_scene(self.ctx, cp, len(prims), C.byref(w)))
        self._scene_key = key
        return prims, world

    def set_grid_copies(self, copies: int):
        """Batch-tally copies (power of two) for grids bound from now on."""
        N.check(self.lib.cc_set_grid_copies(self.ctx, int(copies)))
        self._grid_key = None

    def atomic_rate(self, n_slots: int, n_ops: int):
        """FP64 atomic adds per second at random addresses of an n_slots-double array (the bound
        of the voxel tallies) and the array total after the timed and warm-up launches."""
        rate, total = C.c_double(), C.c_double()
        N.check(self.lib.cc_atomic_rate(self.ctx, int(n_slots), int(n_ops), C.byref(rate), C.byref(total)))
        return rate.value, total.value

    def bind_grid(self, origin, shape, voxel_cm, track_uncertainty):
        key = (tuple(float(x) for x in origin), tuple(int(x) for x in shape), float(voxel_cm),
               bool(track_uncertainty))
        if key != self._grid_key:
            o = np.ascontiguousarray(origin, dtype=np.float64)
            s = np.ascontiguousarray(shape, dtype=np.int32)
            N.check(self.lib.cc_bind_grid(self.ctx, N.ptr(o, C.c_double), N.ptr(s, C.c_int32),
                                          float(voxel_cm), 1 if track_uncertainty else 0))
            self._grid_key = key
        else:
            N.check(self.lib.cc_reset_grid(self.ctx))

    def read_grid(self, shape, moments=True, out=None):
        """Totals (and with moments the batch moments and hit counts) of the bound grid.  out:
        optional dict of C-contiguous arrays of the grid's shape (kerma, h10, kerma_sum2, h10_sum2
        float64, hits int32), filled in place instead of new arrays, e.g. a VoxelGrid's own."""
        nv = int(np.prod(shape))
        out = out or {}

        def target(key, dtype):
            a = out.get(key)
            if a is None:
                return np.empty(nv, dtype=dtype)
            if a.dtype != dtype or a.size != nv or not a.flags.c_contiguous:
                raise ValueError(f"read_grid: {key} must be a C-contiguous {dtype} array of {nv} voxels")
            return a.reshape(nv)
        k = target("kerma", np.float64)
        h = target("h10", np.float64)
        q1 = target("kerma_sum2", np.float64) if moments else None
        q2 = target("h10_sum2", np.float64) if moments else None
        hits = target("hits", np.int32) if moments else None
        nb = C.c_int64()
        nh = C.c_int64()
        nul = C.POINTER(C.c_double)()
        N.check(self.lib.cc_read_grid(
            self.ctx, N.ptr(k, C.c_double), N.ptr(h, C.c_double),
            N.ptr(q1, C.c_double) if moments else nul, N.ptr(q2, C.c_double) if moments else nul,
            N.ptr(hits, C.c_int32) if moments else C.POINTER(C.c_int32)(), C.byref(nb), C.byref(nh)))
        r = dict(kerma=k.reshape(shape), h10=h.reshape(shape), n_batches=nb.value, n_histories=nh.value)
        if moments:
            r.update(kerma_sum2=q1.reshape(shape), h10_sum2=q2.reshape(shape),
                     hits=hits.reshape(shape))
        return r

    def load_grid(self, saved: dict):
        """Adds a grid read by read_grid (totals, and batch moments and hit counts when present)
        into the bound grid, with its batch and history counts: the resume step of a run split
        over history ranges, and the add-into-totals step of a merge (cc_load_grid)."""
        def arr(key, dtype, ctype):
            a = saved.get(key)
            if a is None:
                return C.POINTER(ctype)()
            a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
            keep.append(a)
            return N.ptr(a, ctype)
        keep = []
        N.check(self.lib.cc_load_grid(
            self.ctx, arr("kerma", np.float64, C.c_double), arr("h10", np.float64, C.c_double),
            arr("kerma_sum2", np.float64, C.c_double), arr("h10_sum2", np.float64, C.c_double),
            arr("hits", np.int32, C.c_int32), int(saved.get("n_batches", 0)), int(saved.get("n_histories", 0))))

    # -- transport ---------------------------------------------------------------
    def run(self, src: SourceRecord, n, seed, first=0, batch=200_000, flags=0) -> N.CcRunOut:
        out = N.CcRunOut()
        N.check(self.lib.cc_run(self.ctx, C.byref(src.struct), int(n), int(seed), int(first),
                                int(batch), int(flags), C.byref(out)))
        return out

    def transport(self, pos, dirv, energy, seed, first=0, flags=0, trace=False):
        """cc_transport on copies of the inputs; with trace=True also returns the segment log
        (TRACE_DTYPE records; the call is repeated with a larger buffer if the first overflowed,
        which gives the same results since they depend on the seed only)."""
        n = pos.shape[0]
        cap = max(4096, 64 * n)
        while True:
            p = np.array(pos, dtype=np.float64, order="C")
            d = np.array(dirv, dtype=np.float64, order="C")
            e = np.array(energy, dtype=np.float64, order="C")
            nsc = np.zeros(n, dtype=np.int32)
            ab = np.zeros(n, dtype=np.uint8)
            es = np.zeros(n, dtype=np.uint8)
            out = N.CcRunOut()
            f = int(flags)
            if trace:
                N.check(self.lib.cc_trace_alloc(self.ctx, cap))
                f |= N.RUN_TRACE
            N.check(self.lib.cc_transport(self.ctx, n, N.ptr(p, C.c_double), N.ptr(d, C.c_double),
                                          N.ptr(e, C.c_double), int(seed), int(first), f,
                                          N.ptr(nsc, C.c_int32), N.ptr(ab, C.c_uint8), N.ptr(es, C.c_uint8),
                                          C.byref(out)))
            if not trace:
                return p, d, e, nsc, ab.astype(bool), es.astype(bool), out
            total = C.c_uint64()
            N.check(self.lib.cc_trace_read(self.ctx, 0, None, C.byref(total)))
            if total.value <= cap:
                rec = np.zeros(total.value, dtype=N.TRACE_DTYPE)
                N.check(self.lib.cc_trace_read(self.ctx, total.value, rec.ctypes.data_as(C.c_void_p),
                                               C.byref(total)))
                return p, d, e, nsc, ab.astype(bool), es.astype(bool), out, rec
            cap = int(total.value)
            if flags & N.RUN_GRID:
                N.check(self.lib.cc_reset_grid(self.ctx))   # the repeated call scores the grid again

    # -- queries -------------------------------------------------------------------
    def locate(self, geometry, pts):
        self.set_geometry(geometry)
        q = _f32(pts)
        m = np.zeros(len(q), dtype=np.int32)
        N.check(self.lib.cc_locate(self.ctx, len(q), N.ptr(q, C.c_float), N.ptr(m, C.c_int32)))
        return m

    def boundary(self, geometry, o, d, eps=1e-6):
        self.set_geometry(geometry)
        oo, dd = _f32(o), _f32(d)
        t = np.zeros(len(oo), dtype=np.float32)
        esc = np.zeros(len(oo), dtype=np.int32)
        N.check(self.lib.cc_test_boundary(self.ctx, len(oo), N.ptr(oo, C.c_float), N.ptr(dd, C.c_float),
                                          float(eps), N.ptr(t, C.c_float), N.ptr(esc, C.c_int32)))
        return t, esc.astype(bool), None


_engine = None
_engine_lock = threading.Lock()


def default_engine() -> Engine:
    global _engine
    with _engine_lock:
        if _engine is None:
            _engine = Engine(0)
        return _engine


def _seed_from(seed):
    if seed is None:
        return int(np.random.SeedSequence().generate_state(2, dtype=np.uint32).view(np.uint64)[0])
    return int(seed) & 0xFFFFFFFFFFFFFFFF


def transport_photons(pos: np.ndarray, dirv: np.ndarray, energy: np.ndarray, geometry: Geometry,
                      rng: np.random.Generator, grid: VoxelGrid | None = None, recorder=None,
                      tally_rng: np.random.Generator | None = None,
                      fluorescence_enabled: bool = True) -> BatchResult:
    """chatcarlo/transport.py:98-233 on the GPU; pos/dirv/energy are updated in place.

    The Philox seed is drawn from `rng` (advancing it, as the reference's
    transport does); `tally_rng` is accepted for signature compatibility and
    unused because the DDA tally draws no random numbers.  `recorder` (a
    trajectory.TrajectoryRecorder) receives every flight segment from the GPU's
    segment log; logging draws no random numbers, so the results are the same
    with and without it.
    """
    eng = default_engine()
    eng.set_geometry(geometry)
    flags = N.RUN_FLUORESCENCE if fluorescence_enabled else 0
    if grid is not None:
        eng.bind_grid(grid.origin_cm, grid.shape, grid.voxel_size_cm, False)
        flags |= N.RUN_GRID
    seed = int(rng.integers(0, 2 ** 63))
    if recorder is not None:
        from .trajectory import fill_recorder
        p, d, e, nsc, ab, es, out, rec = eng.transport(pos, dirv, energy, seed, 0, flags, trace=True)
        fill_recorder(recorder, rec, eng.pack.materials)
    else:
        p, d, e, nsc, ab, es, out = eng.transport(pos, dirv, energy, seed, 0, flags)
    pos[...] = p
    dirv[...] = d
    energy[...] = e
    if grid is not None:
        g = eng.read_grid(grid.shape, moments=False)
        grid.kerma_keV += g["kerma"]
        grid.h10_track_pSv_cm3 += g["h10"]
    names = eng.pack.materials
    edep = {names[m]: float(out.edep_keV[m]) for m in range(len(names)) if out.edep_keV[m] != 0.0}
    return BatchResult(n_scatter=nsc.astype(int), absorbed=ab, escaped=es, final_energy=energy,
                       energy_deposited=edep, n_fluorescence=int(out.n_fluorescence))


def run_transport(scene, n_histories: int = 100_000, seed: int | None = None,
                  batch_size: int = 200_000, dose_grid: bool = False,
                  grid_resolution_cm: float = 5.0, n_workers: int = 1,
                  track_uncertainty: bool = True) -> TransportResult:
    """chatcarlo/transport.py:358-474 on one GPU."""
    src = scene.raw["source"]
    geometry = Geometry(scene.raw["geometry"])
    eng = default_engine()
    eng.set_geometry(geometry)
    grid = (VoxelGrid.from_bbox(geometry.bbox_min, geometry.bbox_max, grid_resolution_cm,
                                track_uncertainty=track_uncertainty) if dose_grid else None)
    fluor = scene.raw.get("physics", {}).get("fluorescence", True)
    flags = N.RUN_FLUORESCENCE if fluor else 0
    if grid is not None:
        eng.bind_grid(grid.origin_cm, grid.shape, grid.voxel_size_cm, track_uncertainty)
        flags |= N.RUN_GRID
        if track_uncertainty:
            flags |= N.RUN_TRACK_UNCERTAINTY
    n_mas = None
    if src.get("mas") is not None and src.get("ctdi_vol_mGy") is None:
        # before the run, so that a calibration that cannot be made (no SpekPy) fails at once
        from .source import photon_count_through_field
        n_mas = photon_count_through_field(src)
    rec = SourceRecord(src)
    out = eng.run(rec, n_histories, _seed_from(seed), 0, batch_size, flags)
    names = eng.pack.materials
    used = [m for m in range(len(names)) if out.edep_keV[m] != 0.0]
    edep_keV = {names[m]: float(out.edep_keV[m]) for m in used}
    moments = ScalarMoments()
    if track_uncertainty:
        moments.sums = dict(edep_keV)
        moments.sum2s = {names[m]: float(out.edep_sum2[m]) for m in used}
        moments.n_batches = int(out.n_batches)
        moments.n_histories = int(n_histories)
    if grid is not None:
        # straight into the grid's own arrays (no intermediate copies)
        dst = {"kerma": grid.kerma_keV, "h10": grid.h10_track_pSv_cm3}
        if track_uncertainty:
            dst.update(kerma_sum2=grid.kerma_sum2, h10_sum2=grid.h10_sum2, hits=grid.n_batches_hit)
        g = eng.read_grid(grid.shape, moments=track_uncertainty, out=dst)
        if track_uncertainty:
            grid.n_batches = int(g["n_batches"])
            grid.n_histories = int(g["n_histories"])
    n_real = n_mas   # chatcarlo/transport.py:448-449
    if src.get("ctdi_vol_mGy") is not None:
        # after the grid is read: the calibration binds the phantom's own scene and grid
        # (chatcarlo/transport.py:445-447, same master seed)
        from .ctdi import effective_histories_from_ctdi
        n_real = effective_histories_from_ctdi(src, seed=seed, engine=eng)
    n = max(int(n_histories), 1)
    if track_uncertainty:
        sem = {k: v / 1000.0 for k, v in moments.standard_errors().items()}
        rel = moments.relative_errors()
        n_batches = moments.n_batches
    else:
        # the reference's count when statistics are off: the grid's batch count, which end_batch
        # leaves at 0 then (chatcarlo/transport.py:455-459, tally.py:59-64)
        sem, rel = {}, {}
        n_batches = grid.n_batches if grid is not None else 0
    stats = dict(n_events=int(out.n_events), n_flights=int(out.n_flights),
                 n_voxel_visits=int(out.n_voxel_visits), rounds_max=int(out.rounds_max),
                 kernel_ms=list(out.kernel_ms), kernel_launches=list(out.kernel_launches))
    return TransportResult(
        n_histories=int(n_histories),
        energy_deposited_MeV={k: v / 1000.0 for k, v in edep_keV.items()},
        fraction_absorbed=out.n_absorbed / n, fraction_escaped=out.n_escaped / n,
        mean_scatter_events=out.scatter_sum / n, grid=grid, n_photons_real=n_real,
        n_fluorescence=int(out.n_fluorescence), n_batches=n_batches,
        energy_deposited_sem_MeV=sem, energy_deposited_rel_err=rel, device_stats=stats)
