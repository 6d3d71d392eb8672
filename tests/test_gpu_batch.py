"""Batched transport (many scenes fed by one device work queue, in two forms: one persistent
history-based launch, and event-based rounds in one particle pool replayed as a hipGraph).

Every scene's history h uses the Philox stream keyed by (seed_s, h), exactly as
history h of a single-scene run with seed_s, so a batched run must reproduce,
scene by scene, the kerma and H*(10) grids and the deposits of the single-scene
runs up to the order of FP64 additions, and the event counters exactly.  The
event-based launch shapes below include pools far smaller than the history count, so
slots are recycled through the free-slot stack many times and the queue crosses scene
boundaries inside one refill; the persistent ones flush the LDS tally cache from every
iteration to about once per workgroup.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes(n_scenes=5, seed=11):
    from chatcarlo_amd.batch import SceneSpec
    from chatcarlo_amd.workloads import batched_scenes
    rng = np.random.default_rng(seed)
    out = []
    for k, (raw, grid, _, s) in enumerate(batched_scenes(n_scenes, master_seed=2026 + seed)):
        vox = [10.0, 6.0, 15.0, 10.0, 12.0][k % 5]
        shape = tuple(int(300 // vox) for _ in range(3))
        out.append(SceneSpec(raw, (np.array([-150.0, -150.0, -150.0]), shape, vox),
                             int(rng.integers(2000, 20000)), s))
    return out


@pytest.fixture(scope="module")
def single(engine):
    """Per-scene single-scene runs: grids, deposits, counters."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    scenes = _scenes()
    res = []
    for s in scenes:
        engine.set_geometry(Geometry(s.raw["geometry"]))
        engine.bind_grid(s.grid[0], s.grid[1], s.grid[2], False)
        out = engine.run(SourceRecord(s.raw["source"]), s.n_histories, s.seed, 0, 4096,
                         N.RUN_FLUORESCENCE | N.RUN_GRID)
        g = engine.read_grid(s.grid[1], moments=False)
        res.append(dict(kerma=g["kerma"], h10=g["h10"], edep=np.array(out.edep_keV[:]),
                        cnt=np.array([out.n_absorbed, out.n_escaped, out.scatter_sum, out.n_fluorescence,
                                      out.n_events, out.n_flights, out.n_voxel_visits], dtype=np.int64)))
    return scenes, res


@pytest.mark.parametrize("tune", [(256, 8, 1 << 20), (128, 4, 1024), (64, 2, 4096), (192, 16, 3000)])
def test_batched_equals_single_scene_runs(engine, single, tune):
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.batch import BatchedRun, Tuning
    scenes, ref = single
    b = BatchedRun(scenes, engine)
    st = b.run(Tuning(*tune))
    print(f"tuning {tune}: {st.n_histories} histories, {st.rounds} rounds, {st.device_ms:.2f} ms", flush=True)
    assert st.n_histories == sum(s.n_histories for s in scenes)
    got = np.array([st.n_absorbed, st.n_escaped, st.scatter_sum, st.n_fluorescence, st.n_events, st.n_flights,
                    st.n_voxel_visits], dtype=np.int64)
    np.testing.assert_array_equal(got, np.sum([r["cnt"] for r in ref], axis=0))
    assert st.n_absorbed + st.n_escaped == st.n_histories
    names = engine.pack.materials
    for k, r in enumerate(ref):
        kk, hh, edep = b.read(k)
        np.testing.assert_allclose(kk, r["kerma"], rtol=1e-9, atol=1e-12 * r["kerma"].max())
        np.testing.assert_allclose(hh, r["h10"], rtol=1e-9, atol=1e-12 * r["h10"].max())
        e = np.zeros(N.CC_MAX_MAT)
        for m, v in edep.items():
            e[names.index(m)] = v
        np.testing.assert_allclose(e, r["edep"], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("block,epoch", [(512, 2), (256, 1), (512, 1000), (256, 7)])
def test_persistent_equals_single_scene_runs(engine, single, block, epoch):
    """The one persistent launch: the same counters as the single-scene runs, the same grids
    and deposits up to FP64 summation order.  epoch 1 flushes the LDS tally cache after every
    iteration, epoch 1000 about once per workgroup."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.batch import BatchedRun, Tuning
    scenes, ref = single
    b = BatchedRun(scenes, engine)
    st = b.run(Tuning(block, tally_copies=4, persistent=True, epoch=epoch))
    print(f"persistent block {block} epoch {epoch}: {st.n_histories} histories, {st.device_ms:.2f} ms", flush=True)
    assert st.graph_launches == 1 and st.rounds == 0
    got = np.array([st.n_absorbed, st.n_escaped, st.scatter_sum, st.n_fluorescence, st.n_events, st.n_flights,
                    st.n_voxel_visits], dtype=np.int64)
    np.testing.assert_array_equal(got, np.sum([r["cnt"] for r in ref], axis=0))
    names = engine.pack.materials
    for k, r in enumerate(ref):
        kk, hh, edep = b.read(k)
        np.testing.assert_allclose(kk, r["kerma"], rtol=1e-9, atol=1e-12 * r["kerma"].max())
        np.testing.assert_allclose(hh, r["h10"], rtol=1e-9, atol=1e-12 * r["h10"].max())
        e = np.zeros(N.CC_MAX_MAT)
        for m, v in edep.items():
            e[names.index(m)] = v
        np.testing.assert_allclose(e, r["edep"], rtol=1e-9, atol=1e-9)


def test_persistent_with_a_few_histories_per_scene(engine):
    """Scenes of 1, 2 and 3 histories: most lanes never get a history, and a wave's claim spans
    scene boundaries.  The counters equal the single-scene runs' and the grids agree."""
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.batch import BatchedRun, SceneSpec, Tuning
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    base = _scenes(3, seed=17)
    scenes = [SceneSpec(s.raw, s.grid, k + 1, s.seed) for k, s in enumerate(base)]
    ref = np.zeros(7, dtype=np.int64)
    grids = []
    for s in scenes:
        engine.set_geometry(Geometry(s.raw["geometry"]))
        engine.bind_grid(s.grid[0], s.grid[1], s.grid[2], False)
        o = engine.run(SourceRecord(s.raw["source"]), s.n_histories, s.seed, 0, 4096, N.RUN_FLUORESCENCE | N.RUN_GRID)
        ref += np.array([o.n_absorbed, o.n_escaped, o.scatter_sum, o.n_fluorescence, o.n_events, o.n_flights,
                         o.n_voxel_visits], dtype=np.int64)
        grids.append(engine.read_grid(s.grid[1], moments=False)["kerma"])
    b = BatchedRun(scenes, engine)
    for t in (Tuning(512, tally_copies=2, persistent=True, epoch=2), Tuning(256, tally_copies=1, persistent=True, epoch=1)):
        st = b.run(t)
        assert st.n_histories == 6
        got = np.array([st.n_absorbed, st.n_escaped, st.scatter_sum, st.n_fluorescence, st.n_events, st.n_flights,
                        st.n_voxel_visits], dtype=np.int64)
        np.testing.assert_array_equal(got, ref)
        for k, gk in enumerate(grids):
            np.testing.assert_allclose(b.read(k)[0], gk, rtol=1e-9, atol=1e-12 * max(gk.max(), 1e-30))


def test_fraction_and_rerun_start_from_zero(engine, single):
    from chatcarlo_amd.batch import BatchedRun, Tuning
    scenes, ref = single
    b = BatchedRun(scenes, engine)
    st = b.run(Tuning(256, 4, 8192), fraction=0.25)
    assert st.n_histories == sum(max(1, int(np.ceil(s.n_histories * 0.25))) for s in scenes)
    st = b.run(Tuning(256, 4, 8192))
    for k, r in enumerate(ref):
        kk, _, _ = b.read(k)
        np.testing.assert_allclose(kk, r["kerma"], rtol=1e-9, atol=1e-12 * r["kerma"].max())
    res = b.results()
    assert [x.n_histories for x in res] == [s.n_histories for s in scenes]
    np.testing.assert_allclose(res[0].grid.kerma_keV, ref[0]["kerma"], rtol=1e-9, atol=1e-12 * ref[0]["kerma"].max())


def test_autotune_returns_a_candidate(engine, single):
    from chatcarlo_amd.batch import BatchedRun, Tuning
    b = BatchedRun(_scenes(3, seed=5), engine)
    t, table = b.autotune(fraction=0.2, cached=False)
    # the persistent form at blocks 512 and 256, then the event-based form
    assert table.shape == (3, 7) and np.all(table[:, 6] > 0)
    assert [int(x) for x in table[:, 0]] == [1, 1, 0]
    assert [int(x) for x in table[:2, 1]] == [512, 256] and [int(x) for x in table[:, 2]] == [2, 2, 2]
    assert (int(table[2, 1]), int(table[2, 3]), int(table[2, 4])) == (128, 4, 1 << 19)
    # the first run's shape unless a later one beats the best so far by 5 %
    best = table[0]
    for row in table[1:]:
        if row[6] > 1.05 * best[6]:
            best = row
    assert (int(t.persistent), t.block, t.epoch, t.rounds_per_sync, t.pool, t.tally_copies) == \
        tuple(int(x) for x in best[:6])
    # the same workload again: the cached shape, no timing
    t2, table2 = b.autotune(fraction=0.2)
    assert t2 == t and np.array_equal(table2, table)
    # the timing runs leave no tallies behind: a run after tuning equals the single-scene runs
    scenes, ref = single
    b = BatchedRun(scenes, engine)
    b.autotune(fraction=0.1, cached=False)
    b.run(Tuning(128, 4, 8192))
    for k, r in enumerate(ref):
        kk, _, _ = b.read(k)
        np.testing.assert_allclose(kk, r["kerma"], rtol=1e-9, atol=1e-12 * r["kerma"].max())


def test_bad_tuning_is_rejected(engine):
    from chatcarlo_amd.batch import BatchedRun, Tuning
    b = BatchedRun(_scenes(2, seed=3), engine)
    for t in (Tuning(100, 8, 4096), Tuning(256, 3, 4096), Tuning(256, 8, 10), Tuning(512, 8, 4096),
              Tuning(256, persistent=True, epoch=0), Tuning(320, persistent=True), Tuning(128, persistent=True)):
        with pytest.raises(ValueError):
            b.run(t)
    with pytest.raises(ValueError):
        b.run(Tuning(), fraction=0.0)


def _config4_picks():
    """Three scenes of the 64-scene sweep: the one with the most primitives, one with lead (the
    most lead primitives among the rest) and the one with the highest kVp among the rest."""
    from chatcarlo_amd.workloads import batched_scenes
    allsc = batched_scenes(64)
    nprim = [len(x[0]["geometry"]) - 1 for x in allsc]
    kvp = [max(b["energy_keV"] for b in x[0]["source"]["spectrum"]) for x in allsc]
    lead = [sum(g["material"] == "lead" for g in x[0]["geometry"]) for x in allsc]
    a = int(np.argmax(nprim))
    b = max((i for i in range(64) if i != a), key=lambda i: (lead[i], nprim[i]))
    c = max((i for i in range(64) if i not in (a, b)), key=lambda i: kvp[i])
    assert lead[b] > 0 and kvp[c] > 145.0
    return [allsc[i] for i in (a, b, c)]


def test_sweep_scenes_on_their_2cm_grids_chi2_against_oracle(engine, dpack):
    """The 64-scene sweep's scenes at their real grids (2 cm voxels, 150^3 over the 300 cm room):
    three of them run batched, and kerma and H*(10) are compared voxel by voxel with the FP64
    oracle by chi^2.  Batched runs keep totals only, so the GPU's variance comes from eight
    independent repetitions (other seeds); the oracle tallies the scene grid's voxels inside the
    box that holds the GPU's well-converged voxels, over several processes."""
    import sys
    sys.path.insert(0, __import__("os").path.dirname(__file__))
    from test_gpu_transport import _box_around, _chi2_voxels
    from chatcarlo_amd.batch import BatchedRun, SceneSpec, Tuning
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.source import SourceRecord
    from oracle import core
    from oracle.parallel import run_parallel
    picks = _config4_picks()
    reps, n_gpu = 8, 1_000_000
    tot = [[[], []] for _ in picks]
    for r in range(reps):
        specs = [SceneSpec(raw, grid, n_gpu, (seed + 1_000_003 * r) % (1 << 62)) for raw, grid, _, seed in picks]
        b = BatchedRun(specs, engine)
        st = b.run(Tuning(128, 4, 1 << 20))
        assert st.n_histories == n_gpu * len(picks)
        for k in range(len(picks)):
            kk, hh, _ = b.read(k)
            tot[k][0].append(kk / n_gpu)
            tot[k][1].append(hh / n_gpu)
    for k, (raw, (origin, shape, vox), _, seed) in enumerate(picks):
        geo = Geometry(raw["geometry"])
        rec = SourceRecord(raw["source"])
        gk = np.array(tot[k][0])
        m1k = gk.mean(axis=0)
        v1k = gk.var(axis=0, ddof=1) / reps
        good = (m1k > 0) & (np.sqrt(v1k) < 0.01 * m1k)
        lo, hi = _box_around(good, (64, 64, 64))
        sub = tuple(slice(int(a), int(c)) for a, c in zip(lo, hi))
        prims, world = geo.packed(dpack.mat_id)
        n_or = 400_000
        print(f"scene {k}: {len(raw['geometry']) - 1} primitives, box {lo}..{hi}, oracle {n_or}", flush=True)
        _, grid = run_parallel(raw["source"], prims, world, dpack, rec.spec_e, rec.spec_cdf, n_or, seed % (1 << 32),
                               2_000, grid_spec=(np.asarray(origin) + lo * vox, tuple(int(x) for x in hi - lo), vox))
        for key, q, i in (("kerma", "kerma_q", 0), ("h10", "h10_q", 1)):
            g = np.array(tot[k][i])[(slice(None),) + sub]
            m1, v1 = g.mean(axis=0), g.var(axis=0, ddof=1) / reps
            m2 = getattr(grid, key) / grid.n_histories
            v2 = core.batch_variance(getattr(grid, key), getattr(grid, q), grid.n_batches, grid.n_histories) / grid.n_histories
            _chi2_voxels(m1, v1, m2, v2, 200, f"sweep scene {k} {key}")
