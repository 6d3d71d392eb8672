"""CPU-side checks: the C-ABI library, host logic, data-pack invariants, oracle transport physics."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import core
from oracle import transport as OT

ROOT = Path(__file__).resolve().parent.parent
RNG = np.random.default_rng(11)


# ---------------------------------------------------------------------------
# C-ABI
# ---------------------------------------------------------------------------
def _header_functions():
    text = (ROOT / "include" / "chatcarlo_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(cc_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from chatcarlo_amd import _native as N
    lib = N.load()
    declared = _header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(N.EXPORTS) == declared


def test_ctypes_layouts_match_the_header(tmp_path):
    from chatcarlo_amd import _native as N
    checks = [("sizeof(cc_prim)", C.sizeof(N.CcPrim)), ("sizeof(cc_world)", C.sizeof(N.CcWorld)),
              ("sizeof(cc_tables)", C.sizeof(N.CcTables)), ("sizeof(cc_source)", C.sizeof(N.CcSource)),
              ("offsetof(cc_source, heel_span)", N.CcSource.heel_span.offset),
              ("offsetof(cc_source, heel_cdf)", N.CcSource.heel_cdf.offset),
              ("sizeof(cc_run_out)", C.sizeof(N.CcRunOut)),
              ("offsetof(cc_run_out, edep_keV)", N.CcRunOut.edep_keV.offset),
              ("offsetof(cc_run_out, region_sum)", N.CcRunOut.region_sum.offset),
              ("offsetof(cc_run_out, n_records)", N.CcRunOut.n_records.offset),
              ("sizeof(cc_train_cfg)", C.sizeof(N.CcTrainCfg)),
              ("offsetof(cc_train_cfg, seed)", N.CcTrainCfg.seed.offset),
              ("offsetof(cc_train_cfg, target_power)", N.CcTrainCfg.target_power.offset),
              ("offsetof(cc_importance_cfg, split_gain)", N.CcImportanceCfg.split_gain.offset),
              ("sizeof(cc_train_out)", C.sizeof(N.CcTrainOut)),
              ("offsetof(cc_train_out, n_positive)", N.CcTrainOut.n_positive.offset),
              ("sizeof(cc_importance_cfg)", C.sizeof(N.CcImportanceCfg)),
              ("offsetof(cc_importance_cfg, seed)", N.CcImportanceCfg.seed.offset),
              ("sizeof(cc_trace_record)", C.sizeof(N.CcTraceRecord)),
              ("sizeof(cc_trace_record)", N.TRACE_DTYPE.itemsize),
              ("offsetof(cc_trace_record, event)", N.CcTraceRecord.event.offset),
              ("offsetof(cc_trace_record, event)", N.TRACE_DTYPE.fields["event"][1]),
              ("CC_RUN_TRACE", N.RUN_TRACE), ("CC_EV_FLUORESCENCE", N.EV_FLUORESCENCE),
              ("CC_RUN_PERSISTENT", N.RUN_PERSISTENT),
              ("CC_NOALLOC_MAX_PHOTONS", N.NOALLOC_MAX_PHOTONS),
              ("CC_EV_RAYLEIGH", N.EV_RAYLEIGH),
              ("sizeof(cc_batch_tuning)", C.sizeof(N.CcBatchTuning)),
              ("offsetof(cc_batch_tuning, pool)", N.CcBatchTuning.pool.offset),
              ("offsetof(cc_batch_tuning, epoch)", N.CcBatchTuning.epoch.offset),
              ("sizeof(cc_batch_out)", C.sizeof(N.CcBatchOut)),
              ("offsetof(cc_batch_out, device_ms)", N.CcBatchOut.device_ms.offset),
              ("offsetof(cc_batch_out, error_bits)", N.CcBatchOut.error_bits.offset),
              ("CC_MLP_PARAMS", N.MLP_PARAMS), ("CC_RECORD_FLOATS", N.RECORD_FLOATS),
              ("CC_RUN_SOURCE_BIAS", N.RUN_SOURCE_BIAS), ("CC_RUN_IMPORTANCE", N.RUN_IMPORTANCE),
              ("CC_RUN_RECORD", N.RUN_RECORD), ("CC_RUN_REGION", N.RUN_REGION)]
    body = "".join(f'printf("%zu\\n", (size_t)({expr}));' for expr, _ in checks)
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "chatcarlo_hip.h"\n'
                   f'int main(void){{{body}return 0;}}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [v for _, v in checks]
    from chatcarlo_amd.geometry import PRIM_DTYPE, WORLD_DTYPE
    assert PRIM_DTYPE.itemsize == C.sizeof(N.CcPrim) == 48
    assert WORLD_DTYPE.itemsize == C.sizeof(N.CcWorld)


def test_missing_library_fails_loudly(tmp_path):
    from chatcarlo_amd import _native as N
    with pytest.raises(N.NativeError):
        N.load(tmp_path / "nope.so")


# ---------------------------------------------------------------------------
# scene validation (reference tests/test_scene.py style)
# ---------------------------------------------------------------------------
def _scene(**over):
    raw = {"source": {"kvp": 80, "position": [0, 0, 100], "direction": [0, 0, -2],
                      "field": {"size_cm": [30, 30], "sid_cm": 100}},
           "geometry": [{"shape": "box", "material": "water", "center": [0, 0, 0], "size_cm": [30, 30, 20]}]}
    raw.update(over)
    return raw


def test_scene_valid_and_normalised():
    from chatcarlo_amd.scene import validate_scene
    sc = validate_scene(_scene())
    assert sc.ok, sc.errors
    assert sc.raw["source"]["direction"] == [0.0, 0.0, -1.0]
    assert sc.raw["source"]["field"]["shape"] == "rect"
    assert sc.raw["geometry"][0]["name"] == "object_0"


@pytest.mark.parametrize("mutate,path", [
    (lambda r: r["source"].update(kvp=200), "source.kvp"),
    (lambda r: r["source"].update(spectrum=[{"energy_keV": 160, "weight": 1}], kvp=None), "source.spectrum"),
    (lambda r: r["source"].update(direction=[0, 0, 0]), "source.direction"),
    (lambda r: r["source"]["field"].update(shape="hexagon"), "source.field.shape"),
    (lambda r: r["source"]["field"].pop("sid_cm"), "source.field.sid_cm"),
    (lambda r: r["geometry"][0].update(shape="torus"), "geometry[0].shape"),
    (lambda r: r["geometry"][0].update(size_cm=[1, -1, 1]), "geometry[0].size_cm"),
    (lambda r: r.update(physics={"fluorescence": "yes"}), "physics.fluorescence"),
    (lambda r: r["source"].update(rotation={"isocenter": [0, 0, 0], "n_angles": 1}), "source.rotation.n_angles"),
])
def test_scene_errors_have_paths(mutate, path):
    from chatcarlo_amd.scene import validate_scene
    raw = _scene()
    mutate(raw)
    if raw["source"].get("kvp") is None:
        raw["source"].pop("kvp", None)
    sc = validate_scene(raw)
    assert not sc.ok
    assert any(e.path == path for e in sc.errors), [str(e) for e in sc.errors]


def test_unknown_material_raises_value_error():
    from chatcarlo_amd.materials import mu_rho
    with pytest.raises(ValueError, match="candidates"):
        mu_rho("unobtanium", 60.0)


def test_world_box_and_grid_shape():
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.tally import VoxelGrid
    from chatcarlo_amd.workloads import chest_room
    raw, _ = chest_room()
    g = Geometry(raw["geometry"])
    np.testing.assert_allclose(g.bbox_min, [-270, -270, -70])
    np.testing.assert_allclose(g.bbox_max, [270, 270, 370])
    vg = VoxelGrid.from_bbox(g.bbox_min, g.bbox_max, 5.0)
    assert vg.shape == (108, 108, 88)


# ---------------------------------------------------------------------------
# data pack invariants
# ---------------------------------------------------------------------------
def test_pack_invariants(pack):
    assert np.all(np.diff(pack.e_node) > 0)
    assert pack.e_node[0] == 1.0 and pack.e_node[-1] == 150.0
    assert np.all(pack.e_node.astype(np.float32).astype(np.float64) == pack.e_node)
    assert np.all(pack.sig > 0)
    np.testing.assert_allclose(pack.sig[:, 4], pack.sig[:, 0] + pack.sig[:, 1] + pack.sig[:, 2], rtol=1e-12)
    for m, cdf in enumerate(pack.sel_cdf):
        assert np.all(np.diff(cdf, axis=2) >= -1e-12)
        np.testing.assert_allclose(cdf[:, :, -1], 1.0)
    assert np.all((pack.sq_over_z >= 0) & (pack.sq_over_z <= 1.0 + 1e-12))
    assert np.all(np.diff(pack.ray_a, axis=1) >= 0)
    # F(Z,0) = Z (reference tests/test_rayleigh.py:20-29): A(x) ~ Z^2 x near 0
    for i, z in enumerate(pack.elements):
        assert pack.ray_a[i, 1] / pack.ray_x[1] == pytest.approx(z * z, rel=5e-3)
    # S(Z,q) -> Z for large q (reference tests/test_bound_compton.py:17-27)
    assert np.all(pack.sq_over_z[:, -1] > 0.98)


# ---------------------------------------------------------------------------
# batch statistics (reference tests/test_uncertainty.py:22-50, 83-121)
# ---------------------------------------------------------------------------
def test_batch_statistics_brute_force_at_batch_size_one():
    from chatcarlo_amd.tally import relative_error, standard_error
    x = RNG.exponential(2.0, 500)
    T, Q = x.sum(), np.sum(x ** 2)
    sem = standard_error(T, Q, len(x), len(x))
    assert sem == pytest.approx(x.std(ddof=1) / np.sqrt(len(x)), rel=1e-12)
    assert relative_error(T, Q, len(x), len(x)) == pytest.approx(sem / x.mean(), rel=1e-12)
    assert np.isnan(standard_error(T, Q, 1, len(x)))
    np.testing.assert_allclose(core.standard_error(T, Q, len(x), len(x)), sem, rtol=1e-12)


def test_scalar_moments_merge_equals_single_stream():
    from chatcarlo_amd.tally import ScalarMoments
    a, b, c = ScalarMoments(), ScalarMoments(), ScalarMoments()
    batches = [({"water": 3.0}, 10), ({"water": 1.0, "bone": 2.0}, 12), ({"bone": 4.0}, 9)]
    for v, n in batches:
        c.add_batch(v, n)
    for v, n in batches[:2]:
        a.add_batch(v, n)
    b.add_batch(*batches[2])
    a.merge_from(b.as_moments(), b.n_batches, b.n_histories)
    assert a.sums == c.sums
    for k in c.sum2s:
        assert a.sum2s[k] == pytest.approx(c.sum2s[k], rel=1e-15)


# ---------------------------------------------------------------------------
# oracle tallies and transport physics (reference tests/test_tally.py, test_transport.py)
# ---------------------------------------------------------------------------
def test_overlap_tally_conservation_and_chords():
    t = np.zeros((10, 10, 10))
    n = 500
    o = np.tile([0.5, 0.5, 0.0], (n, 1))
    d = np.tile([0.0, 0.0, 1.0], (n, 1))
    core.overlap_tally(t, [0, 0, 0], 1.0, o, d, np.full(n, 7.3), np.full(n, 12.0))
    assert t.sum() == pytest.approx(n * 7.3 * 12.0, rel=1e-12)
    assert t[0, 0, 7] == pytest.approx(n * 0.3 * 12.0, rel=1e-12)
    t2 = np.zeros((10, 10, 10))
    o2 = np.tile([-5.0, 5.0, 5.0], (200, 1))
    core.overlap_tally(t2, [0, 0, 0], 1.0, o2, np.tile([1.0, 0, 0], (200, 1)), np.full(200, 20.0), np.full(200, 5.0))
    assert t2.sum() == pytest.approx(200 * 5.0 * 10.0, rel=1e-12)
    # random segments: totals equal the clipped lengths
    m = 2000
    o3 = RNG.uniform(-3, 13, (m, 3))
    d3 = RNG.normal(size=(m, 3))
    d3 /= np.linalg.norm(d3, axis=1, keepdims=True)
    L = RNG.exponential(5.0, m)
    t3 = np.zeros((10, 10, 10))
    core.overlap_tally(t3, [0, 0, 0], 1.0, o3, d3, L, np.ones(m))
    # brute force: fine sampling of each segment
    tot = 0.0
    for i in range(m):
        s = np.linspace(0, L[i], 4001)
        p = o3[i] + d3[i] * s[:, None]
        inside = np.all((p >= 0) & (p <= 10), axis=1)
        tot += np.trapezoid(inside.astype(float), s)
    assert t3.sum() == pytest.approx(tot, rel=2e-3)


def test_stratified_estimator_expectation_is_the_overlap():
    """Reference tests/test_tally.py:115-138: surface voxel of boundary-start segments is unbiased."""
    shape = (20, 10, 10)
    n = 100_000
    L = RNG.exponential(4.86, n).clip(max=19.9)
    o = np.tile([0.0, 5.5, 5.5], (n, 1))
    d = np.tile([1.0, 0.0, 0.0], (n, 1))
    ts = np.zeros(shape)
    core.stratified_tally(ts, [0, 0, 0], 1.0, o, d, L, np.ones(n), np.random.default_rng(12))
    te = np.zeros(shape)
    core.overlap_tally(te, [0, 0, 0], 1.0, o, d, L, np.ones(n))
    exact = np.minimum(L, 1.0).sum()
    assert te[0, 5, 5] == pytest.approx(exact, rel=1e-10)
    assert abs(ts[0, 5, 5] - exact) / exact < 0.005


def _oracle_slab(pack, material, thickness, size=100.0):
    from chatcarlo_amd.geometry import Geometry
    g = Geometry([{"shape": "box", "material": material, "center": [0, 0, 0], "size_cm": [thickness, size, size]}])
    prims, world = g.packed(pack.mat_id)
    return core.OGeom.from_packed(prims, world), g


def test_oracle_beer_lambert(dpack):
    og, _ = _oracle_slab(dpack, "water", 10.0)
    n = 40_000
    pos = np.tile([-15.0, 0, 0], (n, 1))
    d = np.tile([1.0, 0, 0], (n, 1))
    r = OT.transport_photons(pos, d, np.full(n, 60.0), og, dpack, np.random.default_rng(1))
    unc = np.mean(r["escaped"] & (r["n_scatter"] == 0))
    mu = core.xs_lookup(dpack, np.array([dpack.mat_id("water"), dpack.mat_id("air")]), np.array([60.0, 60.0]))["mu"]
    exp = np.exp(-mu[0] * 10 - mu[1] * 60)
    assert abs(unc - exp) < 5 * np.sqrt(exp * (1 - exp) / n)


@pytest.mark.parametrize("fluor", [True, False])
def test_oracle_energy_conservation(dpack, fluor):
    og, _ = _oracle_slab(dpack, "lead", 0.5, 50.0)
    n = 4000
    pos = np.tile([-5.25, 0, 0], (n, 1))
    d = np.tile([1.0, 0, 0], (n, 1))
    e = np.full(n, 100.0)
    r = OT.transport_photons(pos, d, e, og, dpack, np.random.default_rng(42), fluorescence=fluor)
    dep = sum(r["energy_deposited"].values())
    esc = r["final_energy"][r["escaped"]].sum()
    assert abs(dep + esc - 100.0 * n) / (100.0 * n) < 1e-9
    assert (r["n_fluorescence"] > 0) == fluor


def test_oracle_collision_vs_tracklength(dpack):
    og, g = _oracle_slab(dpack, "water", 15.0, 60.0)
    grid = OT.Grid(g.bbox_min, np.ceil((g.bbox_max - g.bbox_min) / 3.0).astype(int), 3.0)
    n = 30_000
    pos = np.tile([-12.5, 0, 0], (n, 1))
    d = np.tile([1.0, 0, 0], (n, 1))
    r = OT.transport_photons(pos, d, np.full(n, 60.0), og, dpack, np.random.default_rng(7), grid=grid)
    coll = sum(r["energy_deposited"].values())
    assert abs(grid.kerma.sum() - coll) / coll < 0.05
