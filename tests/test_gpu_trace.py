"""GPU segment log behind transport_photons(recorder=...): the invariants of chatcarlo's
own trajectory tests (tests/test_trace.py:56-122) on the same slab and beams, plus the
log's agreement with the transport counters."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _slab():
    from chatcarlo_amd.geometry import Geometry
    return Geometry([{"name": "slab", "shape": "box", "material": "water", "center": [0.0, 0.0, 0.0],
                      "size_cm": [10.0, 100.0, 100.0]}])


def _beam(n, e):
    return (np.tile([-30.0, 0.0, 0.0], (n, 1)), np.tile([1.0, 0.0, 0.0], (n, 1)), np.full(n, float(e)))


def _run(n, seed, e=60.0, recorder=None, geom=None, fluor=True):
    from chatcarlo_amd.transport import transport_photons
    pos, dirv, en = _beam(n, e)
    r = transport_photons(pos, dirv, en, geom or _slab(), np.random.default_rng(seed), recorder=recorder,
                          fluorescence_enabled=fluor)
    return r, pos, dirv


def _segments(rec, pid):
    out = []
    for ids, s, e in zip(rec.photon_ids, rec.starts, rec.ends):
        m = ids == pid
        if np.any(m):
            out.append((s[m][0], e[m][0]))
    return out


def test_segments_are_continuous():
    from chatcarlo_amd.trajectory import TrajectoryRecorder
    rec = TrajectoryRecorder()
    _run(50, 1, recorder=rec)
    for pid in range(50):
        segs = _segments(rec, pid)
        assert len(segs) >= 2
        for (_, end_prev), (start_next, _) in zip(segs, segs[1:]):
            assert np.allclose(start_next, end_prev, atol=1e-5)


def test_energy_changes_only_at_compton_and_fluorescence():
    from chatcarlo_amd.trajectory import TrajectoryRecorder, trajectories_to_json
    rec = TrajectoryRecorder()
    _run(500, 2, e=80.0, recorder=rec)
    trajs = trajectories_to_json(rec)
    assert len(trajs) == 500
    saw_drop = saw_same = False
    for t in trajs:
        en, ev = t["energies"], t["events"]
        for i in range(len(en) - 1):
            assert en[i + 1] <= en[i] + 1e-9
            if ev[i] == "compton":
                assert en[i + 1] < en[i]
                saw_drop = True
            elif ev[i] in ("boundary", "rayleigh"):
                assert abs(en[i + 1] - en[i]) < 1e-9
                saw_same = True
    assert saw_drop and saw_same


def test_last_event_is_terminal():
    from chatcarlo_amd.trajectory import TrajectoryRecorder, trajectories_to_json
    rec = TrajectoryRecorder()
    _run(500, 3, recorder=rec)
    for t in trajectories_to_json(rec):
        assert t["events"][-1] in ("photoelectric", "escape")
        assert "photoelectric" not in t["events"][:-1]
        assert "escape" not in t["events"][:-1]


def test_recorder_does_not_change_results():
    from chatcarlo_amd.trajectory import TrajectoryRecorder
    a, pa, da = _run(300, 7)
    rec = TrajectoryRecorder()
    b, pb, db = _run(300, 7, recorder=rec)
    for k in ("absorbed", "escaped", "n_scatter", "final_energy"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert np.array_equal(pa, pb) and np.array_equal(da, db)
    assert a.energy_deposited == b.energy_deposited
    assert rec.photon_ids


def test_log_matches_transport_counters_and_fluorescence():
    """Lead at 100 keV: every segment is logged once (segments = flight steps), the final
    event of each photon agrees with its absorbed/escaped flag, interactions per photon
    equal n_scatter, and the fluorescence relabels equal the fluorescence count."""
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.trajectory import TrajectoryRecorder, trajectories_to_json
    lead = Geometry([{"name": "slab", "shape": "box", "material": "lead", "center": [0.0, 0.0, 0.0],
                      "size_cm": [0.5, 50.0, 50.0]}])
    rec = TrajectoryRecorder()
    r, _, _ = _run(2000, 11, e=100.0, recorder=rec, geom=lead)
    trajs = trajectories_to_json(rec)
    assert len(trajs) == 2000
    order = np.concatenate(rec.photon_ids)
    first_seen = list(dict.fromkeys(order.tolist()))
    nfl = 0
    for pid, t in zip(first_seen, trajs):
        ev = t["events"]
        assert (ev[-1] == "photoelectric") == bool(r.absorbed[pid])
        assert (ev[-1] == "escape") == bool(r.escaped[pid])
        assert sum(e in ("photoelectric", "compton", "rayleigh", "fluorescence") for e in ev) == r.n_scatter[pid]
        nfl += ev.count("fluorescence")
    assert nfl == r.n_fluorescence > 0
    assert {m for ms in rec.materials for m in ms.tolist()} <= {"air", "lead"}


def test_small_buffer_reports_overflow(engine):
    import ctypes as C
    from chatcarlo_amd import _native as N
    engine.set_geometry(_slab())
    pos, dirv, en = _beam(1000, 60.0)
    N.check(engine.lib.cc_trace_alloc(engine.ctx, 100))
    p, d, e = pos.copy(), dirv.copy(), en.copy()   # cc_transport writes the final states back
    out = N.CcRunOut()
    N.check(engine.lib.cc_transport(engine.ctx, 1000, N.ptr(p, C.c_double), N.ptr(d, C.c_double),
                                    N.ptr(e, C.c_double), 5, 0, N.RUN_FLUORESCENCE | N.RUN_TRACE,
                                    None, None, None, C.byref(out)))
    total = C.c_uint64()
    rec = np.zeros(100, dtype=N.TRACE_DTYPE)
    N.check(engine.lib.cc_trace_read(engine.ctx, 100, rec.ctypes.data_as(C.c_void_p), C.byref(total)))
    assert total.value == out.n_events > 100
    assert np.all(rec["photon"] < 1000) and np.all(rec["event"] <= N.EV_RAYLEIGH)
    # the engine repeats the call with a buffer that fits
    got = engine.transport(pos, dirv, en, 5, 0, N.RUN_FLUORESCENCE, trace=True)[-1]
    assert len(got) == out.n_events


def test_trace_cli_writes_json(tmp_path):
    import json

    import yaml
    from chatcarlo_amd.__main__ import main
    from chatcarlo_amd.workloads import _plain, water_slab_60kev
    scene = tmp_path / "slab.yaml"
    scene.write_text(yaml.safe_dump(_plain(water_slab_60kev()[0])))
    out = tmp_path / "t.json"
    assert main(["trace", str(scene), "-n", "40", "--seed", "3", "--out", str(out)]) == 0
    doc = json.loads(out.read_text())
    assert len(doc["trajectories"]) == 40 and sum(doc["categories"].values()) == 40


def _trace_stats(ids_steps, n):
    """Per-history statistics of a segment log given as a list of per-step (photon ids, energies,
    event codes): segments per history, segments ending in each event per history, and the
    energy of the k-th segment (k < 6) averaged over the histories that have one."""
    from chatcarlo_amd.trajectory import EVENT_NAMES
    seg = np.zeros(n)
    per_ev = np.zeros((len(EVENT_NAMES), n))
    e_k = [[] for _ in range(6)]
    for k, (ids, en, ev) in enumerate(ids_steps):
        np.add.at(seg, ids, 1)
        for j in range(len(EVENT_NAMES)):
            np.add.at(per_ev[j], ids[ev == j], 1)
        if k < 6:
            e_k[k].append(np.asarray(en, dtype=float))
    rows = [("segments per history", seg)]
    rows += [(f"{EVENT_NAMES[j]} per history", per_ev[j]) for j in range(len(EVENT_NAMES))]
    stats = [(name, float(x.mean()), float(x.std(ddof=1)) / np.sqrt(n)) for name, x in rows]
    for k in range(6):
        x = np.concatenate(e_k[k]) if e_k[k] else np.zeros(0)
        if len(x) > 100:
            stats.append((f"energy of segment {k}", float(x.mean()), float(x.std(ddof=1)) / np.sqrt(len(x))))
    return stats


def test_segment_log_matches_the_oracle_transport(dpack):
    """The GPU segment log against the oracle's round loop on the same slab and beam (80 keV pencil
    on 10 cm of water): segments per history, the share of segments ending at a boundary, an
    escape, photoabsorption, fluorescence, Compton and Rayleigh events, and the mean energy of
    the first six segments, each within statistics (|z| < 4, sum of z^2 within 3 sigma)."""
    from chatcarlo_amd.trajectory import EVENT_NAMES, TrajectoryRecorder
    from oracle import core
    from oracle import transport as OT
    n = 40_000
    rec = TrajectoryRecorder()
    _run(n, 9, e=80.0, recorder=rec)
    ev_index = {name: j for j, name in enumerate(EVENT_NAMES)}
    gpu = [(ids, en, np.array([ev_index[x] for x in ev])) for ids, en, ev in zip(rec.photon_ids, rec.energies,
                                                                              rec.events)]
    prims, world = _slab().packed(dpack.mat_id)
    og = core.OGeom.from_packed(prims, world)
    steps = []
    pos, dirv, en = _beam(n, 80.0)
    OT.transport_photons(pos, dirv, en, og, dpack, np.random.default_rng(5),
                         recorder=lambda ids, s, e, energy, ev, m: steps.append((ids, energy, ev)))
    a, b = _trace_stats(gpu, n), _trace_stats(steps, n)
    assert [x[0] for x in a] == [x[0] for x in b]
    zs = []
    for (name, m1, s1), (_, m2, s2) in zip(a, b):
        z = (m1 - m2) / np.hypot(s1, s2) if s1 + s2 > 0 else 0.0
        print(f"{name}: gpu {m1:.5g} oracle {m2:.5g} z={z:+.2f}")
        zs.append(z)
    zs = np.array(zs)
    assert np.all(np.abs(zs) < 4.0), zs
    assert np.sum(zs ** 2) < len(zs) + 3 * np.sqrt(2 * len(zs))
    assert dict((x[0], x[1]) for x in a)["fluorescence per history"] == 0.0   # water: no K lines above 5 keV
