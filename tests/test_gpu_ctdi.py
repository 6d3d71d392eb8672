"""CTDIvol calibration on the GPU: chatcarlo's own checks (tests/test_ctdi.py: finite and
positive, head phantom above body, periphery above centre in the body phantom, effective photon
count x CTDIw = CTDIvol), run_transport's n_photons_real, and the CLI's calibrated outputs."""
import functools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_SRC = {"kvp": 120.0, "filtration_mm_al": 2.5, "position": [0.0, 0.0, 170.0], "direction": [0.0, 0.0, -1.0],
        "rotation": {"isocenter": [0.0, 0.0, 110.0], "axis": "y", "scan_length_cm": 30.0},
        "field": {"shape": "rect", "size_cm": [50.0, 4.0], "sid_cm": 60.0}}


@functools.lru_cache(maxsize=None)
def _ctdi(phantom):
    from chatcarlo_amd.ctdi import ctdi_per_history_Gy
    return ctdi_per_history_Gy(_SRC, phantom=phantom, n_histories=400_000, seed=31)


def test_ctdi_positive_and_finite():
    for v in _ctdi("body"):
        assert np.isfinite(v) and v > 0


def test_head_above_body_and_periphery_above_centre():
    assert _ctdi("head")[0] > _ctdi("body")[0]
    _, dc, dp = _ctdi("body")
    assert dp > dc


def test_effective_histories_consistency():
    from chatcarlo_amd.ctdi import ctdi_per_history_Gy, effective_histories_from_ctdi
    src = {**_SRC, "ctdi_vol_mGy": 12.0, "ctdi_phantom": "body"}
    n_eff = effective_histories_from_ctdi(src, seed=31)
    ctdiw, _, _ = ctdi_per_history_Gy(src, phantom="body", n_histories=1_000_000, seed=77)
    assert n_eff * ctdiw * 1e3 == pytest.approx(12.0, rel=0.1)


def test_run_transport_calibrates_and_cli_writes_absolute_maps(tmp_path):
    import yaml
    from chatcarlo_amd.__main__ import main
    from chatcarlo_amd.scene import validate_scene
    from chatcarlo_amd.transport import run_transport
    raw = {"source": {**_SRC, "ctdi_vol_mGy": 10.0},
           "geometry": [{"name": "patient", "shape": "cylinder", "material": "water", "center": [0, 0, 110],
                         "radius_cm": 15.0, "height_cm": 60.0, "axis": "y"}]}
    sc = validate_scene(raw)
    assert sc.ok, sc.errors
    r = run_transport(sc, n_histories=100_000, seed=5, dose_grid=True, grid_resolution_cm=5.0)
    ctdiw, _, _ = _ctdi("body")
    assert r.n_photons_real == pytest.approx(10.0e-3 / ctdiw, rel=0.1)
    scene = tmp_path / "ct.yaml"
    scene.write_text(yaml.safe_dump(raw))
    out = tmp_path / "d.npz"
    assert main(["run", str(scene), "-n", "50000", "--seed", "3", "--dose-grid", "--dose-out", str(out)]) == 0
    z = np.load(out)
    assert "dose_Gy" in z and "h10_pSv" in z
    scale = z["dose_Gy"].max() / z["dose_per_history_Gy"].max()
    assert scale == pytest.approx(r.n_photons_real, rel=0.15)


def test_ctdi_matches_the_oracle_procedure(dpack):
    """CTDIw, D_center and D_periphery per history on the GPU against the same procedure on the
    FP64 oracle's transport (the phantom, the 1 cm grid, the hole and ring voxels of ctdi.py),
    within statistics: eight GPU runs of 4e5 histories and the oracle over 16 independent
    processes give the spreads; the oracle's dose uses its own point location and the pack's
    densities."""
    from chatcarlo_amd.ctdi import (CHAMBER_LENGTH_CM, GRID_RESOLUTION_CM, PHANTOM_DIAMETER_CM,
                                    PHANTOM_LENGTH_CM, ctdi_from_dose, ctdi_per_history_Gy,
                                    ctdi_phantom_geometry)
    from chatcarlo_amd.source import SourceRecord
    from chatcarlo_amd.tally import VoxelGrid
    from oracle import core
    from oracle.parallel import run_parallel
    assert CHAMBER_LENGTH_CM == 10.0
    gpu = np.array([ctdi_per_history_Gy(_SRC, phantom="body", n_histories=400_000, seed=900 + k)
                    for k in range(8)])
    rot = _SRC["rotation"]
    geo = ctdi_phantom_geometry(rot, "body")
    radius = PHANTOM_DIAMETER_CM["body"] / 2.0
    iso = np.asarray(rot["isocenter"], dtype=float)
    ax = 1
    lo, hi = iso - radius - 2.0, iso + radius + 2.0
    lo[ax] = iso[ax] - PHANTOM_LENGTH_CM / 2.0 - 2.0
    hi[ax] = iso[ax] + PHANTOM_LENGTH_CM / 2.0 + 2.0
    grid = VoxelGrid.from_bbox(lo, hi, GRID_RESOLUTION_CM)
    prims, world = geo.packed(dpack.mat_id)
    og = core.OGeom.from_packed(prims, world)
    rho = dpack.density[og.material_at(grid.voxel_centers())].reshape(grid.shape)
    rec = SourceRecord(_SRC)
    procs = 16
    _, _, parts = run_parallel(_SRC, prims, world, dpack, rec.spec_e, rec.spec_cdf, procs * 25_000, 4321, 5_000,
                               grid_spec=(grid.origin_cm, grid.shape, grid.voxel_size_cm), processes=procs,
                               per_process=True)
    orc = []
    for p in parts:
        dose = p["grid"]["kerma"] * 1.602176634e-13 / (rho * grid.voxel_volume_cm3()) / p["grid"]["n_histories"]
        orc.append(ctdi_from_dose(dose, grid, iso, ax, radius))
    orc = np.array(orc)
    for j, name in enumerate(("CTDIw", "D_center", "D_periphery")):
        m1, s1 = gpu[:, j].mean(), gpu[:, j].std(ddof=1) / np.sqrt(len(gpu))
        m2, s2 = orc[:, j].mean(), orc[:, j].std(ddof=1) / np.sqrt(len(orc))
        z = (m1 - m2) / np.hypot(s1, s2)
        print(f"{name}: gpu {m1:.5g} +- {s1:.2g}, oracle {m2:.5g} +- {s2:.2g}, z={z:+.2f}")
        assert abs(z) < 4.0, name
