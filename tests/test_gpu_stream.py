"""Streamed runs (the batches of cc_run fed from a device history cursor, in both forms:
event-based rounds in one pool whose last photons one persistent launch carries to their ends,
and one persistent history-based launch per group of batches) against the batch-by-batch loop
(CC_RUN_PER_BATCH): the same histories on the same Philox streams, so tallies, deposits, batch
moments and counters agree up to FP64 summation order (counters and hit counts exactly).  Small
batches make several groups of batches in flight; runs of at least two pools (2 x 2^20 histories)
run as several lanes on their own streams, each over its own part of a group's histories."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(engine, raw, grid, n, batch, seed, form, first=0):
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    engine.set_geometry(Geometry(raw["geometry"]))
    engine.bind_grid(grid[0], grid[1], grid[2], True)
    flags = N.RUN_FLUORESCENCE | N.RUN_GRID | N.RUN_TRACK_UNCERTAINTY | {
        "rounds": 0, "persistent": N.RUN_PERSISTENT, "per_batch": N.RUN_PER_BATCH}[form]
    out = engine.run(SourceRecord(raw["source"]), n, seed, first, batch, flags)
    g = engine.read_grid(grid[1], moments=True)
    cnt = np.array([out.n_absorbed, out.n_escaped, out.scatter_sum, out.n_fluorescence, out.n_events,
                    out.n_flights, out.n_voxel_visits, out.n_batches], dtype=np.int64)
    return out, g, cnt


@pytest.mark.parametrize("form", ["persistent", "rounds"])
@pytest.mark.parametrize("case", ["slab_many_groups", "chest_room", "ct_helical_offset", "chest_room_lanes"])
def test_streamed_equals_batch_by_batch(engine, case, form):
    from chatcarlo_amd import workloads as W
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.tally import VoxelGrid
    if case == "slab_many_groups":
        raw, grid = W.water_slab_60kev()
        n, batch, first = 150_000, 1000, 0          # 150 batches: groups of 64, 64, 22
    elif case.startswith("chest_room"):
        raw, _ = W.chest_room()
        geo = Geometry(raw["geometry"])
        vg = VoxelGrid.from_bbox(geo.bbox_min, geo.bbox_max, 5.0)
        grid = (vg.origin_cm, vg.shape, vg.voxel_size_cm)
        if case == "chest_room":
            n, batch, first = 1_000_003, 100_000, 0     # a short last batch
        else:   # lanes; 72 batches in groups of 64 and 8, lanes split inside batches
            n, batch, first = 5_000_011, 70_000, 123_456
    else:
        raw, grid = W.ct_helical()
        n, batch, first = 300_000, 30_000, 7_777_777
    a, ga, ca = _run(engine, raw, grid, n, batch, 4242, form, first)
    b, gb, cb = _run(engine, raw, grid, n, batch, 4242, "per_batch", first)
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_allclose(np.array(a.edep_keV[:]), np.array(b.edep_keV[:]), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(np.array(a.edep_sum2[:]), np.array(b.edep_sum2[:]), rtol=1e-9, atol=1e-12)
    for k in ("kerma", "h10", "kerma_sum2", "h10_sum2"):
        scale = float(np.abs(gb[k]).max())
        np.testing.assert_allclose(ga[k], gb[k], rtol=1e-9, atol=1e-13 * scale, err_msg=k)
    np.testing.assert_array_equal(ga["hits"], gb["hits"])
    assert ga["n_batches"] == gb["n_batches"] and ga["n_histories"] == gb["n_histories"] == n


@pytest.mark.parametrize("n", [500_000, 4_500_001])
def test_streamed_run_without_grid(engine, n):
    from chatcarlo_amd import _native as N
    from chatcarlo_amd import workloads as W
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    raw, _ = W.chest_room()
    engine.set_geometry(Geometry(raw["geometry"]))
    rec = SourceRecord(raw["source"])
    b = engine.run(rec, n, 9, 0, 50_000, N.RUN_FLUORESCENCE | N.RUN_PER_BATCH)
    for extra in (0, N.RUN_PERSISTENT):
        a = engine.run(rec, n, 9, 0, 50_000, N.RUN_FLUORESCENCE | extra)
        assert (a.n_absorbed, a.n_escaped, a.scatter_sum, a.n_events) == (b.n_absorbed, b.n_escaped, b.scatter_sum,
                                                                          b.n_events)
        np.testing.assert_allclose(np.array(a.edep_keV[:]), np.array(b.edep_keV[:]), rtol=1e-9, atol=1e-12)


def test_kernel_timing_of_lanes(engine):
    """With CC_RUN_TIMING the run reports per kernel kind the summed launch durations and the time
    during which at least one launch of the kind ran: in lanes the launches overlap (busy time below
    the summed durations, and below the run's wall time); one after the other (the batch-by-batch
    loop) the two are the same."""
    import time
    from chatcarlo_amd import _native as N
    from chatcarlo_amd import workloads as W
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    raw, _ = W.chest_room()
    engine.set_geometry(Geometry(raw["geometry"]))
    rec = SourceRecord(raw["source"])
    flags = N.RUN_FLUORESCENCE | N.RUN_TIMING
    engine.run(rec, 4_500_000, 3, 0, 500_000, flags)
    t0 = time.perf_counter()
    a = engine.run(rec, 4_500_000, 3, 0, 500_000, flags)
    wall_ms = (time.perf_counter() - t0) * 1e3
    ms, busy, launches = list(a.kernel_ms), list(a.kernel_busy_ms), list(a.kernel_launches)
    assert launches[1] > 0 and busy[1] > 0
    for k in range(8):
        assert busy[k] <= ms[k] * 1.001 + 0.01, (k, busy[k], ms[k])
    assert busy[1] < 0.95 * ms[1]         # the lanes' flight launches overlap
    assert busy[1] < wall_ms
    b = engine.run(rec, 500_000, 3, 0, 500_000, flags | N.RUN_PER_BATCH)
    assert list(b.kernel_busy_ms) == list(b.kernel_ms) and b.kernel_launches[1] > 0


@pytest.mark.parametrize("n_split", [100_000, 2_400_000])
def test_checkpoint_and_resume_equal_one_run(engine, n_split):
    """Checkpoint and resume: a run split at a batch boundary into [0, n1) and [n1, n) (the second
    started with first_history = n1 after the first part's grid is read and loaded back into a
    freshly bound grid) equals the one run of [0, n): histories are keyed by index, so both parts
    draw the same streams; totals and moments agree to FP64 summation order, hit counts and batch
    and history counts exactly.  The longer case runs its parts in lanes."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd import workloads as W
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    raw, grid = W.water_slab_60kev()
    engine.set_geometry(Geometry(raw["geometry"]))
    rec = SourceRecord(raw["source"])
    flags = N.RUN_FLUORESCENCE | N.RUN_GRID | N.RUN_TRACK_UNCERTAINTY
    n, batch = 2 * n_split, 20_000
    engine.bind_grid(grid[0], grid[1], grid[2], True)
    one = engine.run(rec, n, 77, 0, batch, flags)
    g1 = engine.read_grid(grid[1], moments=True)
    engine.bind_grid(grid[0], grid[1], grid[2], True)
    a = engine.run(rec, n_split, 77, 0, batch, flags)
    saved = engine.read_grid(grid[1], moments=True)
    engine.bind_grid(grid[0], grid[1], grid[2], True)
    engine.load_grid(saved)
    b = engine.run(rec, n - n_split, 77, n_split, batch, flags)
    g2 = engine.read_grid(grid[1], moments=True)
    assert (a.n_absorbed + b.n_absorbed, a.n_escaped + b.n_escaped, a.scatter_sum + b.scatter_sum,
            a.n_voxel_visits + b.n_voxel_visits) == (one.n_absorbed, one.n_escaped, one.scatter_sum,
                                                       one.n_voxel_visits)
    for k in ("kerma", "h10", "kerma_sum2", "h10_sum2"):
        np.testing.assert_allclose(g2[k], g1[k], rtol=1e-9, atol=1e-13 * float(np.abs(g1[k]).max()), err_msg=k)
    np.testing.assert_array_equal(g2["hits"], g1["hits"])
    assert (g2["n_batches"], g2["n_histories"]) == (g1["n_batches"], g1["n_histories"]) == (n // batch, n)
