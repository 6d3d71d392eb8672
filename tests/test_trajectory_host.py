"""Host side of the trajectory log: grouping GPU segment records into the recorder's
per-step arrays, per-photon polylines, and the fate classification (chatcarlo/trajectory.py:
46-132), on hand-made records."""
import numpy as np

from chatcarlo_amd import _native as N
from chatcarlo_amd.trajectory import TrajectoryRecorder, classify_trajectory, fill_recorder, trajectories_to_json


def _rec(rows):
    r = np.zeros(len(rows), dtype=N.TRACE_DTYPE)
    for k, (pid, fl, st, mat, a, b, e, ev) in enumerate(rows):
        r[k] = (pid, fl, st, mat, a, b, e, ev)
    return r


def test_fill_recorder_orders_segments_in_time():
    B, E, P, F, C, R = (N.EV_BOUNDARY, N.EV_ESCAPE, N.EV_PHOTOELECTRIC, N.EV_FLUORESCENCE, N.EV_COMPTON,
                        N.EV_RAYLEIGH)
    rows = [  # photon 1: boundary, compton, (new flight) escape; photon 0: photoelectric at once
        (1, 1, 0, 0, (5, 0, 0), (9, 1, 0), 50.0, E),
        (0, 0, 0, 1, (0, 0, 0), (2, 0, 0), 60.0, P),
        (1, 0, 1, 1, (1, 0, 0), (5, 0, 0), 70.0, C),
        (1, 0, 0, 0, (0, 0, 0), (1, 0, 0), 70.0, B),
    ]
    rec = TrajectoryRecorder()
    fill_recorder(rec, _rec(rows)[np.random.default_rng(1).permutation(4)], ["air", "water"])
    assert [ids.tolist() for ids in rec.photon_ids] == [[0, 1], [1], [1]]
    assert rec.events[0].tolist() == ["photoelectric", "boundary"]
    assert rec.events[1].tolist() == ["compton"] and rec.events[2].tolist() == ["escape"]
    assert rec.materials[0].tolist() == ["water", "air"]
    trajs = trajectories_to_json(rec)
    t1 = trajs[1]
    assert t1["points"] == [[0, 0, 0], [1, 0, 0], [5, 0, 0], [9, 1, 0]]
    assert t1["energies"] == [70.0, 70.0, 50.0]
    assert t1["summary"] == {"n_interactions": 1, "fate": "escaped", "backscattered": False,
                             "category": "scattered_escape"}
    assert trajs[0]["summary"]["category"] == "absorbed"
    _ = (F, R)


def test_classify_trajectory_categories():
    def t(points, events):
        return {"points": points, "events": events}
    assert classify_trajectory(t([[0, 0, 0], [1, 0, 0], [9, 0, 0]], ["boundary", "escape"]))["category"] == "direct"
    back = classify_trajectory(t([[0, 0, 0], [2, 0, 0], [-5, 1, 0]], ["compton", "escape"]))
    assert back["category"] == "backscatter" and back["backscattered"] and back["n_interactions"] == 1
    fl = classify_trajectory(t([[0, 0, 0], [1, 0, 0], [1, 3, 0], [1, 3, 2]], ["fluorescence", "rayleigh",
                                                                              "photoelectric"]))
    assert fl == {"n_interactions": 3, "fate": "absorbed", "backscattered": False, "category": "absorbed"}


def test_empty_recorder():
    rec = TrajectoryRecorder()
    fill_recorder(rec, np.zeros(0, dtype=N.TRACE_DTYPE), ["air"])
    assert trajectories_to_json(rec) == []
