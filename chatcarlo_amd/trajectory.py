"""Photon trajectories for small runs: the recorder and its per-photon polylines.

Drop-in for chatcarlo/trajectory.py:15-132.  `transport_photons(..., recorder=r)`
fills a TrajectoryRecorder from the GPU's segment log (CC_RUN_TRACE): one
`record` call per transport step, holding the k-th segment of every photon that
has one, as the reference's loop records one segment per photon per iteration
(chatcarlo/transport.py:219-229).  Logging draws no random numbers, so results
with and without a recorder are identical.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

EVENT_NAMES = ("boundary", "escape", "photoelectric", "fluorescence", "compton", "rayleigh")
INTERACTIONS = frozenset(("photoelectric", "compton", "rayleigh", "fluorescence"))


@dataclass
class TrajectoryRecorder:
    """Per-step arrays of flight segments (chatcarlo/trajectory.py:15-43).

    Each `record` call appends one array per field; `event` names what ends the
    segment ("boundary", "photoelectric", "compton", "rayleigh", "fluorescence",
    "escape") and `material` the material the segment runs through."""
    starts: list = field(default_factory=list)
    ends: list = field(default_factory=list)
    energies: list = field(default_factory=list)
    events: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    photon_ids: list = field(default_factory=list)

    def record(self, photon_id, start, end, energy_keV, event, material=None) -> None:
        ids = np.asarray(photon_id)
        self.photon_ids.append(ids)
        self.starts.append(np.asarray(start))
        self.ends.append(np.asarray(end))
        self.energies.append(np.asarray(energy_keV, dtype=float))
        self.events.append(np.asarray(event, dtype=object))
        self.materials.append(np.asarray(material if material is not None else [""] * len(ids), dtype=object))


def fill_recorder(recorder: TrajectoryRecorder, records: np.ndarray, material_names) -> None:
    """Appends a GPU segment log (cc_trace_record array) to `recorder`, one `record` call per
    step k holding the k-th segment (in time order) of every photon that has one."""
    if len(records) == 0:
        return
    order = np.lexsort((records["step"], records["flight"], records["photon"]))
    r = records[order]
    pid = r["photon"].astype(np.int64)
    first = np.r_[0, np.flatnonzero(pid[1:] != pid[:-1]) + 1]
    rank = np.arange(len(r)) - np.repeat(first, np.diff(np.r_[first, len(r)]))
    names = np.asarray(material_names, dtype=object)
    ev = np.asarray(EVENT_NAMES, dtype=object)
    for k in range(int(rank.max()) + 1):
        sel = np.flatnonzero(rank == k)
        sel = sel[np.argsort(pid[sel], kind="stable")]
        recorder.record(pid[sel], r["start"][sel].astype(float), r["end"][sel].astype(float),
                        r["energy_keV"][sel].astype(float), ev[r["event"][sel]], names[r["material"][sel]])


def trajectories_to_json(recorder: TrajectoryRecorder) -> list[dict]:
    """Per-photon polylines in first-seen order (chatcarlo/trajectory.py:46-77): points (the
    first start, then every segment end), energies, events and materials per segment, and
    `summary` from classify_trajectory."""
    if not recorder.photon_ids:
        return []
    ids = np.concatenate(recorder.photon_ids)
    starts = np.concatenate(recorder.starts)
    ends = np.concatenate(recorder.ends)
    energies = np.concatenate(recorder.energies)
    events = np.concatenate(recorder.events)
    materials = np.concatenate(recorder.materials)
    out: dict[int, dict] = {}
    for i, p in enumerate(ids.tolist()):
        t = out.get(p)
        if t is None:
            t = out[p] = {"points": [starts[i].tolist()], "energies": [], "events": [], "materials": []}
        t["points"].append(ends[i].tolist())
        t["energies"].append(float(energies[i]))
        t["events"].append(str(events[i]))
        t["materials"].append(str(materials[i]))
    trajs = list(out.values())
    for t in trajs:
        t["summary"] = classify_trajectory(t)
    return trajs


def classify_trajectory(traj: dict) -> dict:
    """Fate of one photon (chatcarlo/trajectory.py:80-132): interaction count, "absorbed" (last
    event photoelectric) or "escaped", backscatter (first and last segment directions at an
    obtuse angle) and the category direct / scattered_escape / backscatter / absorbed."""
    events = traj["events"]
    pts = np.asarray(traj["points"], dtype=float)
    n_int = sum(1 for e in events if e in INTERACTIONS)
    fate = "absorbed" if events[-1] == "photoelectric" else "escaped"
    a = pts[1] - pts[0]
    b = pts[-1] - pts[-2]
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    back = bool(na > 0 and nb > 0 and float(np.dot(a, b)) / (na * nb) < 0)
    if fate == "absorbed":
        cat = "absorbed"
    elif back:
        cat = "backscatter"
    elif n_int == 0:
        cat = "direct"
    else:
        cat = "scattered_escape"
    return {"n_interactions": n_int, "fate": fate, "backscattered": back, "category": cat}
