"""`python -m chatcarlo_amd run scene.yaml ...`: the reference's `chatcarlo run` on the GPU.

Mirrors chatcarlo/__main__.py:131-253 and the `run` parser at :327-349 (flags
-n/--n-histories, --seed, --dose-grid, --resolution, --dose-out,
--no-uncertainty, --workers) and writes the same npz keys.  `validate` is kept
for the scene front end.  `trace` (chatcarlo/__main__.py:48-85) records the
trajectories of a few photons on the GPU and writes them as JSON (the
reference renders them into an HTML preview; rendering is not part of this
package).
"""
from __future__ import annotations

import argparse
import sys

import numpy as np


def cmd_validate(args) -> int:
    from .scene import load_scene
    scene = load_scene(args.scene)
    for w in scene.warnings:
        print(f"[warning] {w}")
    for e in scene.errors:
        print(f"[error] {e}", file=sys.stderr)
    if scene.ok:
        print("OK")
    return 0 if scene.ok else 1


def cmd_run(args) -> int:
    from .diagnostics import background_medium_warning, dose_map_Gy, max_voxel_position_cm, near_source_air_warning
    from .geometry import Geometry
    from .scene import load_scene
    from .transport import run_transport
    scene = load_scene(args.scene)
    if not scene.ok:
        for e in scene.errors:
            print(f"[error] {e}", file=sys.stderr)
        return 1
    for w in scene.warnings:
        print(f"[warning] {w}")
    result = run_transport(scene, n_histories=int(args.n_histories), seed=args.seed,
                           dose_grid=args.dose_grid, grid_resolution_cm=args.resolution,
                           n_workers=args.workers, track_uncertainty=not args.no_uncertainty,
                           batch_size=args.batch_size)
    print(f"histories: {result.n_histories:,}")
    print(f"fraction absorbed (photoelectric): {result.fraction_absorbed:.4f}")
    print(f"fraction escaped: {result.fraction_escaped:.4f}")
    print(f"mean interactions per photon: {result.mean_scatter_events:.4f}")
    print(f"fluorescence emissions: {result.n_fluorescence:,}")
    print("energy deposited per material [MeV, all histories]:")
    for name, e in sorted(result.energy_deposited_MeV.items(), key=lambda kv: -kv[1]):
        rel = result.energy_deposited_rel_err.get(name)
        print(f"  {name}: {e:.6g}" + (f"  (R = {rel:.3g})" if rel is not None and np.isfinite(rel) else ""))
    src = scene.raw["source"]
    scale = result.n_photons_real
    if scale is not None:
        print(f"dose calibration: CTDIvol = {src['ctdi_vol_mGy']:g} mGy ({src.get('ctdi_phantom', 'body')} "
              f"phantom), effective photon count = {scale:.6g}")
    if args.dose_grid:
        geometry = Geometry(scene.raw["geometry"])
        grid = result.grid
        dose = dose_map_Gy(grid, geometry)
        h10 = grid.h10_map_pSv()
        n = result.n_histories
        dose_ph, h10_ph = dose / n, h10 / n
        print(f"dose grid: shape={grid.shape}, resolution={grid.voxel_size_cm} cm")
        print(f"  max absorbed dose [Gy/history]: {dose_ph.max():.6g}")
        print(f"  total kerma: {grid.total_kerma_MeV():.6g} MeV ({grid.total_kerma_MeV() / n * 1000:.6g} keV/history)")
        print(f"  max H*(10) [pSv/history]: {h10_ph.max():.6g}")
        if scale is not None:
            print(f"  max absorbed dose [Gy] (CTDIvol-calibrated): {dose_ph.max() * scale:.6g}")
            print(f"  max H*(10) [pSv] (CTDIvol-calibrated): {h10_ph.max() * scale:.6g}")
        # is a maximum sitting where it means nothing (chatcarlo/__main__.py:213-226)?
        src_pos = np.asarray(src["position"], dtype=float)
        dpos = max_voxel_position_cm(grid, dose_ph)
        w = background_medium_warning(str(geometry.material_at(dpos[None, :])[0]), geometry.background)
        if w:
            print(f"[warning] {w}")
        hpos = max_voxel_position_cm(grid, h10_ph)
        w = near_source_air_warning(str(geometry.material_at(hpos[None, :])[0]), geometry.background,
                                    float(np.linalg.norm(hpos - src_pos)),
                                    geometry.nearest_object_distance_cm(src_pos))
        if w:
            print(f"[warning] {w}")
        if args.dose_out:
            extra = {} if scale is None else {"dose_Gy": dose_ph * scale, "h10_pSv": h10_ph * scale}
            np.savez(args.dose_out, dose_per_history_Gy=dose_ph, h10_per_history_pSv=h10_ph,
                     kerma_keV=grid.kerma_keV, h10_track_pSv_cm3=grid.h10_track_pSv_cm3,
                     origin_cm=grid.origin_cm, voxel_size_cm=grid.voxel_size_cm, shape=np.array(grid.shape),
                     **extra)
            print(f"wrote dose grid: {args.dose_out}")
    return 0


_TRACE_MAX_N = 2000


def cmd_trace(args) -> int:
    """Trajectories of n source photons (at most 2000, as the reference clamps) as JSON."""
    import json

    from .geometry import Geometry
    from .scene import load_scene
    from .source import sample_source_photons
    from .trajectory import TrajectoryRecorder, trajectories_to_json
    from .transport import transport_photons
    scene = load_scene(args.scene)
    if not scene.ok:
        for e in scene.errors:
            print(f"[error] {e}", file=sys.stderr)
        return 1
    for w in scene.warnings:
        print(f"[warning] {w}")
    n = int(args.n)
    if n > _TRACE_MAX_N:
        print(f"[warning] -n={n} is too large; clamped to {_TRACE_MAX_N}")
        n = _TRACE_MAX_N
    rng = np.random.default_rng(args.seed)
    geometry = Geometry(scene.raw["geometry"])
    pos, dirv, energy = sample_source_photons(scene.raw["source"], n, rng)
    rec = TrajectoryRecorder()
    transport_photons(pos, dirv, energy, geometry, rng, recorder=rec,
                      fluorescence_enabled=scene.raw.get("physics", {}).get("fluorescence", True))
    trajs = trajectories_to_json(rec)
    counts = {}
    for t in trajs:
        c = t["summary"]["category"]
        counts[c] = counts.get(c, 0) + 1
    out = args.out or args.scene.rsplit(".", 1)[0] + "_trace.json"
    with open(out, "w") as f:
        json.dump({"scene": args.scene, "n": n, "seed": args.seed, "categories": counts,
                   "trajectories": trajs}, f)
    print(f"wrote the trajectories of {n} photons to {out}")
    return 0


def main(argv=None) -> int:
    p = argparse.ArgumentParser(prog="chatcarlo_amd")
    sub = p.add_subparsers(dest="cmd", required=True)
    pv = sub.add_parser("validate", help="validate a scene.yaml")
    pv.add_argument("scene")
    pv.set_defaults(func=cmd_validate)
    pr = sub.add_parser("run", help="run photon transport on the GPU")
    pr.add_argument("scene")
    pr.add_argument("-n", "--n-histories", type=float, default=1e5)
    pr.add_argument("--seed", type=int, default=None)
    pr.add_argument("--dose-grid", action="store_true")
    pr.add_argument("--resolution", type=float, default=5.0)
    pr.add_argument("--dose-out")
    pr.add_argument("--no-uncertainty", action="store_true")
    pr.add_argument("--workers", type=int, default=1, help="accepted for compatibility; one GPU is used")
    pr.add_argument("--batch-size", type=int, default=200_000)
    pr.set_defaults(func=cmd_run)
    pt = sub.add_parser("trace", help="record photon trajectories on the GPU (JSON)")
    pt.add_argument("scene")
    pt.add_argument("-n", type=int, default=100)
    pt.add_argument("--seed", type=int, default=None)
    pt.add_argument("--out")
    pt.set_defaults(func=cmd_trace)
    args = p.parse_args(argv)
    return args.func(args)


if __name__ == "__main__":
    sys.exit(main())
