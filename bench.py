"""Benchmark: analog photon transport histories/s on one MI355X (chest room, 80 kVp, 1e7 histories).

One step = one full pass of the transport hot path over the workload's history
count (source sampling, flights with the DDA kerma/H*(10) tally, interactions,
batch moments), inputs resident on the device.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "histories/s; H*(10) FOM 1/(σ²t) at 2% rel err, learned vs analog"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes (roofline accounting)
FLIGHT_STATE_B = 64        # particle state read + write per flight
VOXEL_B = 32               # two FP64 read-modify-write tallies per voxel crossing
EVENT_B = 128              # SoA state read + write per transport event
SOURCE_B = 64              # source write per history


def _profile_figure(name, key):
    """A figure written as `key: value` in a committed rocprofv3 summary under profiles/."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith(f"{key}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_traffic(kernel="k_flight"):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary."""
    return _profile_figure(f"{kernel}_chest_room_hbm_traffic.md", f"{kernel}_traffic_bytes_per_launch")


def measured_atomics(kernel="k_flight"):
    """Atomic requests per launch of the dominant kernel (rocprofv3 TCC_ATOMIC; adds of one instruction
    to one 64-byte segment travel as one request) and requests per add of the atomic-rate calibration
    (random addresses), from the committed summary."""
    name = f"{kernel}_chest_room_atomics.md"
    return (_profile_figure(name, f"{kernel}_atomics_per_launch"),
            _profile_figure(name, "calibration_requests_per_add"))


# atomic-rate calibration: random FP64 adds over the footprint of 16 copies of the chest-room grid
ATOMIC_CAL_SLOTS = 16 * 108 * 108 * 88
ATOMIC_CAL_OPS = 64 << 21


REF_BATCH = 200_000   # run_transport's default batch_size (chatcarlo/transport.py:359)


def _cpu_sample(args):
    """Time the oracle's reference-style NumPy transport on a bounded sample (one process)."""
    raw, n, seed, min_seconds = args
    from chatcarlo_amd.datapack import build_pack
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.spectrum import spectrum_table
    from chatcarlo_amd.tally import VoxelGrid
    from oracle import core
    from oracle import transport as OT
    pack = build_pack()
    geo = Geometry(raw["geometry"])
    prims, world = geo.packed(pack.mat_id)
    og = core.OGeom.from_packed(prims, world)
    vg = VoxelGrid.from_bbox(geo.bbox_min, geo.bbox_max, 5.0)
    grid = OT.Grid(vg.origin_cm, vg.shape, vg.voxel_size_cm)
    e, cdf = spectrum_table(raw["source"])
    done = 0
    t0 = time.perf_counter()
    k = 0
    # min_seconds > 0: stop at n histories or once min_seconds have passed; 0: run all n
    # batches of the reference's default size (chatcarlo/transport.py:359), one OT.run call each
    while done < n and (min_seconds <= 0 or time.perf_counter() - t0 < min_seconds):
        m = min(REF_BATCH, n - done)
        OT.run(raw["source"], og, pack, e, cdf, m, seed + k, batch_size=m, grid=grid, exact_tally=False)
        done += m
        k += 1
    return done, time.perf_counter() - t0


def cpu_baseline(raw, all_cores):
    import multiprocessing as mp
    done, dt = _cpu_sample((raw, 2 * REF_BATCH, 11, 12.0))
    out = {"value": done / dt, "unit": "histories/s", "cores": 1, "kind": "port",
           "sample": f"{done} histories of the same workload in batches of {REF_BATCH} (oracle NumPy "
                     f"restatement of the reference transport, stratified track-length tally, 5 cm grid), "
                     f"one process, {dt:.1f} s"}
    if all_cores:
        nproc = max(1, min(os.cpu_count() or 1, 16))
        ctx = mp.get_context("spawn")
        per = REF_BATCH
        with ctx.Pool(nproc) as pool:
            t0 = time.perf_counter()
            res = pool.map(_cpu_sample, [(raw, per, 100 + i, 0.0) for i in range(nproc)])
            dt2 = time.perf_counter() - t0
        tot = sum(r[0] for r in res)
        slowest = max(r[1] for r in res)
        out["all_cores"] = {"value": tot / slowest, "cores": nproc,
                            "sample": f"{tot} histories over {nproc} processes (one batch of {per} each), "
                                      f"slowest process {slowest:.1f} s of transport, {dt2:.1f} s wall incl. start-up"}
    return out


def _config_grid(raw, grid, voxel_cm=5.0):
    """The workload's own grid, or the reference's default (5 cm over the world box)."""
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.tally import VoxelGrid
    if grid is not None:
        return grid
    geo = Geometry(raw["geometry"])
    vg = VoxelGrid.from_bbox(geo.bbox_min, geo.bbox_max, voxel_cm)
    return vg.origin_cm, vg.shape, vg.voxel_size_cm


def _oracle_counts(raw, grid, n, seed, exact):
    """The oracle's transport of n histories on the workload's grid: (segments, voxel pieces,
    seconds).  exact: exact-overlap tally (voxel pieces = voxel crossings); else the reference's
    stratified sub-step tally (its CPU speed)."""
    from chatcarlo_amd.datapack import build_pack
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.spectrum import spectrum_table
    from oracle import core
    from oracle import transport as OT
    pack = build_pack()
    geo = Geometry(raw["geometry"])
    prims, world = geo.packed(pack.mat_id)
    og = core.OGeom.from_packed(prims, world)
    g = OT.Grid(grid[0], grid[1], grid[2])
    e, cdf = spectrum_table(raw["source"])
    fluor = raw.get("physics", {}).get("fluorescence", True)
    t0 = time.perf_counter()
    r = OT.run(raw["source"], og, pack, e, cdf, n, seed, batch_size=REF_BATCH, grid=g, fluorescence=fluor,
               exact_tally=exact)
    return r["n_segments"], r["n_voxel_pieces"], time.perf_counter() - t0


def _config_cpu_worker(args):
    raw, grid, n, seed = args
    _, _, dt = _oracle_counts(raw, grid, n, seed, False)
    return n, dt


def _gpu_counts(eng, raw, grid, n, timed):
    """The GPU run of n histories on the workload's grid (batches of 10^6, resident grid):
    (events per history, voxel crossings per history, histories/s or None)."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    eng.set_geometry(Geometry(raw["geometry"]))
    flags = N.RUN_GRID | N.RUN_TRACK_UNCERTAINTY | (
        N.RUN_FLUORESCENCE if raw.get("physics", {}).get("fluorescence", True) else 0)
    rec = SourceRecord(raw["source"])
    if timed:
        eng.bind_grid(grid[0], grid[1], grid[2], True)
        eng.run(rec, n // 10, 1, 0, 1_000_000, flags)
    eng.bind_grid(grid[0], grid[1], grid[2], True)
    t0 = time.perf_counter()
    o = eng.run(rec, n, 2, 0, 1_000_000, flags)
    dt = time.perf_counter() - t0
    return o.n_events / n, o.n_voxel_visits / n, (n / dt if timed else None)


def configs_table(eng, n_oracle=200_000, n_gpu=10_000_000, n_scene_oracle=10_000, n_scene_gpu=1_000_000):
    """Per configuration: transport events (flight segments) and voxel crossings per history, from
    the oracle (n_oracle histories, exact-overlap tally) and from the GPU (n_gpu histories); the
    algorithmic bytes per history B_hist = 32 N_vox + 128 N_evt + 64 (from the GPU counts);
    histories/s of the oracle's reference-style transport (stratified tally, one batch of the
    reference's default 200,000 histories) in one process and in one process per host core (up to
    16, one such batch each, rate = all histories / slowest process), and of the GPU (resident
    grid, batches of 10^6).  The batched configuration is given as three of its 64 scenes (the most
    primitives, the most lead, the highest kVp, each on its own 2 cm grid) with rates, and as the
    history-weighted mean counts over all 64 scenes (n_scene_oracle / n_scene_gpu histories each)."""
    import multiprocessing as mp
    from chatcarlo_amd import workloads as W
    cases = [("1_water_slab_60keV_pencil", *W.water_slab_60kev()), ("2_chest_room", *W.chest_room()),
             ("3_ct_helical", *W.ct_helical())]
    sc = W.batched_scenes(64)
    picks = {"most_primitives": max(range(64), key=lambda k: len(sc[k][0]["geometry"])),
             "most_lead": max(range(64), key=lambda k: sum(g["material"] == "lead" for g in sc[k][0]["geometry"])),
             "highest_kvp": max(range(64), key=lambda k: max(b["energy_keV"] for b in sc[k][0]["source"]["spectrum"]))}
    for name, k in picks.items():
        cases.append((f"4_batched_scene_{k}_{name}", sc[k][0], sc[k][1]))
    cases.append(("5_angio_room_analog", *W.angio_room()))
    nproc = max(1, min(os.cpu_count() or 1, 16))
    out = {}
    with mp.get_context("spawn").Pool(nproc) as pool:
        for name, raw, grid in cases:
            grid = _config_grid(raw, grid)
            seg, vox, _ = _oracle_counts(raw, grid, n_oracle, 5, True)
            _, _, dt = _oracle_counts(raw, grid, REF_BATCH, 6, False)
            res = pool.map(_config_cpu_worker, [(raw, grid, REF_BATCH, 100 + i) for i in range(nproc)])
            n_evt, n_vox, rate = _gpu_counts(eng, raw, grid, n_gpu, True)
            out[name] = {"grid": [int(x) for x in grid[1]], "voxel_cm": float(grid[2]),
                         "n_evt_oracle": seg / n_oracle, "n_vox_oracle": vox / n_oracle,
                         "n_evt_gpu": n_evt, "n_vox_gpu": n_vox,
                         "bytes_per_history": VOXEL_B * n_vox + EVENT_B * n_evt + SOURCE_B,
                         "cpu_histories_per_s_1proc": REF_BATCH / dt,
                         "cpu_histories_per_s_all_cores": sum(r[0] for r in res) / max(r[1] for r in res),
                         "gpu_histories_per_s": rate}
            print(f"configs-table: {name} done", file=sys.stderr, flush=True)
    acc = np.zeros(5)
    for k, (raw, grid, n_hist, _) in enumerate(sc):
        seg, vox, _ = _oracle_counts(raw, grid, n_scene_oracle, 5, True)
        n_evt, n_vox, _ = _gpu_counts(eng, raw, grid, n_scene_gpu, False)
        acc += n_hist * np.array([1.0, seg / n_scene_oracle, vox / n_scene_oracle, n_evt, n_vox])
        if k % 16 == 15:
            print(f"configs-table: batched scenes {k + 1} of 64", file=sys.stderr, flush=True)
    m = acc[1:] / acc[0]
    out["4_batched_all_64_scenes_history_weighted"] = {
        "histories": int(acc[0]), "n_evt_oracle": m[0], "n_vox_oracle": m[1], "n_evt_gpu": m[2], "n_vox_gpu": m[3],
        "bytes_per_history": VOXEL_B * m[3] + EVENT_B * m[2] + SOURCE_B}
    return {"oracle_histories": n_oracle, "cpu_timed_histories": REF_BATCH, "cpu_processes_all_cores": nproc,
            "gpu_histories": n_gpu, "configs": out}


def ct_helical_run(eng, n_histories=100_000_000):
    """The CT helical workload at its own size: 1e8 histories of the 120 kVp helical scan on the
    CTDI phantom, kerma and H*(10) on the 0.5 cm grid over the phantom, batch moments, grid read
    back (one warm-up run of 1e7 first)."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    from chatcarlo_amd.workloads import ct_helical
    raw, (origin, shape, vox) = ct_helical()
    eng.set_geometry(Geometry(raw["geometry"]))
    rec = SourceRecord(raw["source"])
    flags = N.RUN_FLUORESCENCE | N.RUN_GRID | N.RUN_TRACK_UNCERTAINTY
    eng.bind_grid(origin, shape, vox, True)
    eng.run(rec, 10_000_000, 5, 0, 1_000_000, flags)
    eng.bind_grid(origin, shape, vox, True)
    t0 = time.perf_counter()
    o = eng.run(rec, n_histories, 6, 0, 1_000_000, flags)
    eng.read_grid(shape, moments=True)
    dt = time.perf_counter() - t0
    return {"workload": "ct_helical_120kVp_6mmAl_pmma_ctdi_phantom_0.5cm_grid", "histories": n_histories,
            "grid": list(shape), "voxel_cm": vox, "seconds": dt, "histories_per_s": n_histories / dt,
            "voxel_visits_per_history": o.n_voxel_visits / n_histories,
            "flights_per_history": o.n_flights / n_histories}


def learned_fom(eng, target=0.02, pairs=5):
    """Angio room, operator sphere: contribution records from analog GPU runs, the MLP trained on
    the GPU, then analog and learned runs to the same relative error of the operator H*(10), for
    `pairs` seed pairs.  The learned run's weights are heavy-tailed, so the relative error of one
    run, and the FOM ratio with it, varies by a factor of a few between seeds; the ratio reported
    is the median over the pairs (each pair listed), the run details those of the median pair."""
    from chatcarlo_amd.importance import LearnedImportance
    from chatcarlo_amd.workloads import angio_room, angio_region
    raw, _ = angio_room()
    c, r = angio_region()
    li = LearnedImportance(eng, raw, c, r)
    t0 = time.perf_counter()
    n_rec, _ = li.generate_records(10_000_000, 32_000_000, keep=0.3, seed=61)
    t1 = time.perf_counter()
    tr = li.train(0, steps=3000, batch=65536, lr=8e-3, lr_end=1e-4)
    t2 = time.perf_counter()
    li.configure()
    runs = []
    for k in range(pairs):
        ana, lrn, chosen, z = li.compare(target_rel_err=target, seed_analog=101 + k, seed_learned=202 + k)
        runs.append((lrn.fom / ana.fom, k, ana, lrn, chosen, z))
    ratio, k, ana, lrn, chosen, z = sorted(runs, key=lambda x: x[0])[pairs // 2]
    return {"workload": "angio_room_90kVp_operator_sphere_r10cm", "target_rel_err": target,
            "seed_pairs": [[101 + j, 202 + j] for j in range(pairs)], "median_pair": k,
            "fom_ratio_per_pair": [x[0] for x in runs], "z_per_pair": [x[5] for x in runs],
            "analog": {"histories": ana.n_histories, "rel_err": ana.rel_err, "seconds": ana.seconds,
                       "fom": ana.fom, "h10_pSv_per_history": ana.mean},
            "learned": {"histories": lrn.n_histories, "rel_err": lrn.rel_err, "seconds": lrn.seconds,
                        "fom": lrn.fom, "h10_pSv_per_history": lrn.mean, "splits": lrn.n_split,
                        "roulette_kills": lrn.n_roulette},
            "fom_ratio": ratio, "z_learned_vs_analog": z, "learned_chosen": chosen is lrn,
            "fom_ratio_with_training": (1.0 / (lrn.rel_err ** 2 * (lrn.seconds + (t1 - t0) + (t2 - t1)))) / ana.fom,
            # the learned run's time to exactly the target (FOM-invariant scaling of its run), the map's
            # fixed cost (records + training) added, against the analog run's own time to the target
            "fom_ratio_with_training_at_target": (1.0 / (target ** 2 * (lrn.seconds * (lrn.rel_err / target) ** 2
                                                                        + (t1 - t0) + (t2 - t1)))) / ana.fom,
            "break_even_runs": ((t1 - t0) + (t2 - t1)) / max(ana.seconds * (ana.rel_err / target) ** 2
                                                           - lrn.seconds * (lrn.rel_err / target) ** 2, 1e-12),
            "records": n_rec, "record_seconds": t1 - t0, "train_seconds": t2 - t1,
            "train_device_ms": tr.ms, "train_loss_first_last": [tr.loss_first, tr.loss_last]}


def batched_sweep(eng, autotune_fraction=0.01, sequential=True):
    """64 random scenes (1-40 primitives each, 50-150 kVp, 1e5-1e7 histories, 2 cm grids) in one
    batched run: form and launch shape autotuned on a fraction of the histories, then the whole
    sweep; the whole sweep again in each form (one persistent launch, replayed event-based rounds);
    beside it the same sweep scene by scene (one single-scene run per scene, as a caller of
    run_transport would loop), whose tallies it must equal."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.batch import BatchedRun, SceneSpec, Tuning
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    from chatcarlo_amd.workloads import batched_scenes
    scenes = [SceneSpec(raw, grid, n, s) for raw, grid, n, s in batched_scenes(64)]
    t0 = time.perf_counter()
    b = BatchedRun(scenes, eng)
    t_up = time.perf_counter() - t0
    b.run(fraction=0.005)   # warm-up of both forms (code objects, pool allocation)
    b.run(Tuning(128, 4, 1 << 19, 16), fraction=0.005)
    t0 = time.perf_counter()
    tune, table = b.autotune(autotune_fraction, cached=False)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    st = b.run(tune)
    wall = time.perf_counter() - t0
    out = {"workload": "64_random_scenes_300cm_air_room_2cm_grids", "scenes": len(scenes),
           "histories": st.n_histories,
           "tuning": {"form": "persistent" if tune.persistent else "event_rounds", "block": tune.block,
                      "epoch": tune.epoch, "rounds_per_sync": tune.rounds_per_sync, "pool": tune.pool,
                      "tally_copies": tune.tally_copies},
           "histories_per_s": st.n_histories / wall, "device_ms": st.device_ms, "wall_s": wall,
           "rounds": st.rounds, "graph_launches": st.graph_launches,
           "voxel_visits_per_history": st.n_voxel_visits / st.n_histories,
           "flights_per_history": st.n_flights / st.n_histories,
           "upload_s": t_up, "autotune_s": t_tune, "autotune_fraction": autotune_fraction,
           "autotune_share_of_sweep": t_tune / wall, "wall_s_with_autotune": wall + t_tune,
           "histories_per_s_with_autotune": st.n_histories / (wall + t_tune),
           "autotune_columns": ["persistent", "block", "epoch", "rounds_per_sync", "pool", "tally_copies",
                                "histories_per_s"],
           "autotune_table": [[int(x) for x in r[:6]] + [round(float(r[6]))] for r in table]}
    # the two forms on the whole sweep: one persistent launch and the replayed event-based rounds
    forms = {}
    for name, t in (("persistent", Tuning(512, tally_copies=16, persistent=True, epoch=2)),
                    ("event_rounds", Tuning(128, 4, 1 << 19, 16))):
        s_ = b.run(t)
        forms[name] = {"histories_per_s_device": s_.n_histories / (s_.device_ms * 1e-3), "device_ms": s_.device_ms,
                       "graph_launches": s_.graph_launches}
    forms["persistent_over_event_rounds"] = (forms["persistent"]["histories_per_s_device"]
                                             / forms["event_rounds"]["histories_per_s_device"])
    out["forms"] = forms
    b.run(tune)   # the tuned run's tallies again, for the comparison below
    if sequential:
        kerma_b = [float(b.read(k)[0].sum()) for k in range(len(scenes))]
        t0 = time.perf_counter()
        kerma_s = []
        for s in scenes:
            eng.set_geometry(Geometry(s.raw["geometry"]))
            eng.bind_grid(s.grid[0], s.grid[1], s.grid[2], False)
            eng.run(SourceRecord(s.raw["source"]), s.n_histories, s.seed, 0, 1_000_000,
                    N.RUN_FLUORESCENCE | N.RUN_GRID)
            kerma_s.append(float(eng.read_grid(s.grid[1], moments=False)["kerma"].sum()))
        seq = time.perf_counter() - t0
        out["sequential"] = {"histories_per_s": st.n_histories / seq, "wall_s": seq,
                             "batched_speedup": seq / wall,
                             "max_rel_diff_total_kerma": max(abs(a - c) / c for a, c in zip(kerma_b, kerma_s) if c)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=float, default=1e7)
    ap.add_argument("--batch", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-all-cores", action="store_true",
                    help="skip the all-cores leg of the CPU baseline (one process per core)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-importance FOM comparison")
    ap.add_argument("--learned-only", action="store_true",
                    help="only the learned-importance comparison (prints its JSON object)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-scene batched sweep")
    ap.add_argument("--no-ct", action="store_true", help="skip the 1e8-history CT helical run")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the 64-scene batched sweep (prints its JSON object)")
    ap.add_argument("--configs-table", action="store_true",
                    help="only the per-configuration events, voxel crossings, bytes and CPU / GPU rates")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (PMC byte-counter calibration)")
    ap.add_argument("--calibrate-atomics", action="store_true",
                    help="only run the FP64 atomic-rate calibration (PMC atomic-counter calibration)")
    args = ap.parse_args()
    if args.calibrate_atomics:
        from chatcarlo_amd.transport import default_engine
        rate, total = default_engine().atomic_rate(ATOMIC_CAL_SLOTS, ATOMIC_CAL_OPS)
        print(json.dumps({"adds_per_s": rate, "adds_per_launch": ATOMIC_CAL_OPS, "launches": 2,
                          "array_total": total}))
        return
    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(histories shard across GPUs only in the design, not in this build)", file=sys.stderr)
        sys.exit(2)

    if args.configs_table:
        from chatcarlo_amd.transport import default_engine
        print(json.dumps(configs_table(default_engine())), flush=True)
        return
    if args.batched_only:
        from chatcarlo_amd.transport import default_engine
        print(json.dumps(batched_sweep(default_engine())), flush=True)
        return
    if args.learned_only:
        from chatcarlo_amd.transport import default_engine
        print(json.dumps(learned_fom(default_engine())), flush=True)
        return

    from chatcarlo_amd import _native as N
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.scene import validate_scene
    from chatcarlo_amd.source import SourceRecord
    from chatcarlo_amd.tally import VoxelGrid
    from chatcarlo_amd.transport import default_engine
    from chatcarlo_amd.workloads import chest_room

    raw, _ = chest_room()
    sc = validate_scene(raw)
    assert sc.ok, sc.errors
    geo = Geometry(sc.raw["geometry"])
    vg = VoxelGrid.from_bbox(geo.bbox_min, geo.bbox_max, 5.0, track_uncertainty=True)
    eng = default_engine()
    eng.set_geometry(geo)
    rec = SourceRecord(sc.raw["source"])
    n = int(args.histories)
    flags = N.RUN_FLUORESCENCE | N.RUN_GRID | N.RUN_TRACK_UNCERTAINTY | N.RUN_TIMING

    def step(k, readback=False):
        # one step: reset the tallies and transport (inputs and tallies resident in HBM); with
        # readback also the grid and its moments copied to the host, as run_transport returns them
        # (chatcarlo/transport.py:441-474) -- the PCIe-inclusive rate, reported beside the value
        eng.bind_grid(vg.origin_cm, vg.shape, vg.voxel_size_cm, True)
        o = eng.run(rec, n, 1000 + k, 0, args.batch, flags)
        if readback:   # into a fresh VoxelGrid, as run_transport returns it
            g = VoxelGrid.from_bbox(geo.bbox_min, geo.bbox_max, 5.0, track_uncertainty=True)
            eng.read_grid(vg.shape, moments=True, out={"kerma": g.kerma_keV, "h10": g.h10_track_pSv_cm3,
                                                       "kerma_sum2": g.kerma_sum2, "h10_sum2": g.h10_sum2,
                                                       "hits": g.n_batches_hit})
        return o

    for k in range(args.warmup):
        step(k)
    outs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        outs.append(step(args.warmup + k))
    dt = time.perf_counter() - t0
    hist = n * args.steps
    value = hist / dt
    # the same step with the grid read back into a fresh VoxelGrid, as run_transport returns it:
    # one untimed step (the process's first reads), then the mean of three
    step(args.warmup + args.steps, readback=True)
    rb_steps = 3
    t0 = time.perf_counter()
    for k in range(rb_steps):
        step(args.warmup + args.steps + 1 + k, readback=True)
    dt_rb = (time.perf_counter() - t0) / rb_steps

    flights = sum(o.n_flights for o in outs)
    visits = sum(o.n_voxel_visits for o in outs)
    events = sum(o.n_events for o in outs)
    # The run's lanes overlap their launches: per kind, kernel_ms sums the launch durations and
    # kernel_busy_ms is the time during which at least one launch of the kind ran.  The flight
    # kernel's rate is its algorithmic bytes over its busy time (over its summed durations, i.e.
    # per launch, when launches do not overlap).
    f_ms = sum(o.kernel_ms[1] for o in outs)
    f_busy_s = sum(o.kernel_busy_ms[1] for o in outs) / 1000.0
    f_launch = sum(o.kernel_launches[1] for o in outs)
    f_bytes = FLIGHT_STATE_B * flights + VOXEL_B * visits
    f_avg_s = f_ms / 1000.0 / max(f_launch, 1)
    achieved = f_bytes / f_busy_s / 1e9
    n_evt = events / hist
    n_vox = visits / hist
    b_hist = VOXEL_B * n_vox + EVENT_B * n_evt + SOURCE_B
    o = outs[-1]
    cal_rate, cal_total = eng.atomic_rate(ATOMIC_CAL_SLOTS, ATOMIC_CAL_OPS)
    atom_launch, req_per_add = measured_atomics("k_flight")
    # atomic requests per second while k_flight runs
    atom_achieved = atom_launch * f_launch / f_busy_s if atom_launch else None
    atom_peak = cal_rate * req_per_add if req_per_add else None
    kernel_share = {name: round(sum(x.kernel_busy_ms[i] for x in outs) / (dt * 1000.0), 4)
                    for i, name in enumerate(["source", "flight", "photo", "compton", "rayleigh", "end_batch", "other",
                                              "interact"])}
    line = {
        "metric": METRIC, "value": value, "unit": "histories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1000.0 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 tallies)", "data": "synthetic",
        "config": {"workload": "chest_room_80kVp_kramers_2.5mmAl_1keV_bins", "histories_per_step": n,
                   "batch": args.batch, "grid": list(vg.shape), "voxel_cm": 5.0,
                   "tallies": "kerma + H*(10) track-length, batch moments", "fluorescence": True},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "kernel_bound": "fp64 atomic request rate beyond the L2s (atomic_roofline), not HBM bandwidth",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_flight"),
                     "kernel": "k_flight",
                     "algorithmic_bytes_per_launch": f_bytes / max(f_launch, 1),
                     "avg_launch_ms": f_avg_s * 1000.0, "launches": int(f_launch),
                     "busy_ms_per_step": f_busy_s * 1000.0 / args.steps,
                     "launch_overlap": f_ms / 1000.0 / f_busy_s,
                     "achieved_per_launch": f_bytes / max(f_launch, 1) / f_avg_s / 1e9,
                     "achieved_from": "algorithmic bytes of every k_flight launch over the time at least one "
                                      "ran (launches of the run's lanes overlap; achieved_per_launch divides "
                                      "by the summed launch durations instead)"},
        "atomic_roofline": {"bound": "fp64 atomic requests (performed beyond the L2s)", "kernel": "k_flight",
                            "achieved": atom_achieved, "peak": atom_peak, "unit": "requests/s",
                            "frac": atom_achieved / atom_peak if atom_achieved and atom_peak else None,
                            "requests_per_launch": atom_launch,
                            "peak_from": f"cc_atomic_rate: {ATOMIC_CAL_OPS} random adds over {ATOMIC_CAL_SLOTS} doubles, "
                                         f"{cal_rate:.4g} adds/s x {req_per_add} requests per add",
                            "calibration_total_ok": cal_total == 2.0 * ATOMIC_CAL_OPS},
        "path": {"voxel_scores_per_s": value * n_vox * 2, "n_evt_per_history": n_evt,
                 "n_vox_per_history": n_vox, "bytes_per_history": b_hist,
                 "path_gbs": value * b_hist / 1e9, "path_frac": value * b_hist / 1e9 / HBM_PEAK_GBS,
                 "kernel_time_share": kernel_share, "rounds_max": int(o.rounds_max)},
        "with_grid_readback": {"histories_per_s": n / dt_rb, "ms_per_step": dt_rb * 1000.0,
                               "readback_bytes": int(np.prod(vg.shape)) * 36, "steps": rb_steps,
                               "note": "PCIe-inclusive: each step reads the totals, moments and hit "
                                       "counts into a fresh VoxelGrid, after one untimed such step"},
        "physics": {"fraction_absorbed": o.n_absorbed / n, "fraction_escaped": o.n_escaped / n,
                    "mean_interactions": o.scatter_sum / n, "n_fluorescence": int(o.n_fluorescence)},
    }
    if not args.no_ct:
        line["ct_helical"] = ct_helical_run(eng)
    if not args.no_batched:
        line["batched_sweep"] = batched_sweep(eng)
    if not args.no_learned:
        line["learned_importance"] = learned_fom(eng)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(raw, not args.no_cpu_all_cores)
        cb = line["cpu_baseline"]
        cb["gpu_over_cpu"] = value / cb["value"]
        if "all_cores" in cb:
            cb["gpu_over_all_cores"] = value / cb["all_cores"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
