"""The oracle over several processes merges its tallies as one run over all the batches."""
import numpy as np

from oracle import core
from oracle import parallel as P


def test_parallel_merge_equals_the_batches_of_its_processes(dpack):
    from chatcarlo_amd.geometry import Geometry
    from chatcarlo_amd.spectrum import spectrum_table
    raw = {"source": {"position": [0.0, 0.0, -15.0], "direction": [0.0, 0.0, 1.0],
                      "field": {"shape": "parallel", "size_cm": [4.0, 4.0]},
                      "spectrum": [{"energy_keV": 60.0, "weight": 1.0}]},
           "geometry": [{"name": "slab", "shape": "box", "material": "water", "center": [0, 0, 0],
                         "size_cm": [10.0, 10.0, 10.0]}]}
    geo = Geometry(raw["geometry"])
    prims, world = geo.packed(dpack.mat_id)
    e, c = spectrum_table(raw["source"])
    spec = (np.array([-5.0, -5.0, -5.0]), (5, 5, 5), 2.0)
    out, grid = P.run_parallel(raw["source"], prims, world, dpack, e, c, 4000, 9, 2000, grid_spec=spec,
                               processes=2)
    assert grid.n_batches == 2 and grid.n_histories == 4000 and out["n_histories"] == 4000
    # the same two shares in this process, from the seeds run_parallel documents
    seeds = np.random.SeedSequence(9).spawn(2)
    parts = [P._worker((raw["source"], prims, world, dpack, e, c, 2000, s, 2000, spec, True)) for s in seeds]
    t = [p["grid"]["kerma"] for p in parts]
    np.testing.assert_allclose(grid.kerma, t[0] + t[1], rtol=1e-12, atol=0)
    np.testing.assert_allclose(grid.kerma_q, (t[0] ** 2 + t[1] ** 2) / 2000, rtol=1e-12, atol=0)
    # two batches of equal size: the batch variance is the two-sample variance of the batch sums
    var = core.batch_variance(grid.kerma, grid.kerma_q, grid.n_batches, grid.n_histories)
    per = np.stack([t[0] / 2000, t[1] / 2000])
    exp = per.var(axis=0, ddof=1) / 2
    np.testing.assert_allclose(var / grid.n_histories, exp, rtol=1e-9, atol=1e-9 * float(exp.max()))
    assert out["n_absorbed"] + out["n_escaped"] == 4000
