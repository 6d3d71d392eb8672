"""The oracle's batch transport over several processes.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The reference splits a run over worker
processes with independent streams and merges their sums in worker order
(chatcarlo/transport.py:315-355, 397-439).  Here each process runs oracle.transport.run on its
own history share with its own stream (SeedSequence(seed).spawn), tallying into its own copy of
the grid; the merge adds the totals and the batch moments (Q = sum over batches of S_b^2 / n_b
and the batch counts add across independent sets of equal-size batches), so the merged
core.batch_variance is that of one run over all the batches.  Processes are spawned, so a caller
that holds a GPU context does not fork it.
"""
from __future__ import annotations

import multiprocessing as mp
import os

import numpy as np

from . import core
from . import transport as OT


def _worker(args):
    src, prims, world, pack, spec_e, spec_cdf, n, seed, batch, grid_spec, fluorescence = args
    og = core.OGeom.from_packed(prims, world)
    grid = OT.Grid(*grid_spec) if grid_spec is not None else None
    r = OT.run(src, og, pack, spec_e, spec_cdf, n, seed, batch_size=batch, grid=grid, fluorescence=fluorescence)
    if grid is not None:
        r["grid"] = dict(kerma=grid.kerma, h10=grid.h10, kerma_q=grid.kerma_q, h10_q=grid.h10_q, hits=grid.hits,
                         n_batches=grid.n_batches, n_histories=grid.n_histories)
    return r


def default_processes() -> int:
    """One process per core, at most 16."""
    return max(1, min(os.cpu_count() or 1, 16))


def run_parallel(src, prims, world, pack, spec_e, spec_cdf, n_histories, seed, batch_size, grid_spec=None,
                 fluorescence=True, processes=None, per_process=False):
    """oracle.transport.run over `processes` processes (each a whole number of batches).

    grid_spec = (origin, shape, voxel) or None.  Returns (run summary, merged Grid or None); the
    summary adds the counters, deposits and region sums of every process.  per_process: also
    returns the list of each process's own result (independent runs, for spreads of derived
    quantities) as a third element."""
    procs = processes or default_processes()
    per = -(-int(n_histories) // procs)
    per = -(-per // batch_size) * batch_size
    shares = []
    left = int(n_histories)
    while left > 0:
        shares.append(min(per, left))
        left -= shares[-1]
    seeds = np.random.SeedSequence(seed).spawn(len(shares))
    jobs = [(src, prims, world, pack, spec_e, spec_cdf, m, s, batch_size, grid_spec, fluorescence)
            for m, s in zip(shares, seeds)]
    if len(jobs) == 1:
        res = [_worker(jobs[0])]
    else:
        with mp.get_context("spawn").Pool(len(jobs)) as pool:
            res = pool.map(_worker, jobs)
    out = {}
    for k in ("n_absorbed", "n_escaped", "scatter_sum", "n_fluorescence", "n_histories", "n_batches",
              "n_segments", "n_voxel_pieces", "region_sum", "region_sum2"):
        out[k] = sum(r[k] for r in res)
    edep = {}
    for r in res:
        for m, v in r["energy_deposited_keV"].items():
            edep[m] = edep.get(m, 0.0) + v
    out["energy_deposited_keV"] = edep
    grid = None
    if grid_spec is not None:
        grid = OT.Grid(*grid_spec)
        for r in res:
            g = r["grid"]
            grid.kerma += g["kerma"]
            grid.h10 += g["h10"]
            grid.kerma_q += g["kerma_q"]
            grid.h10_q += g["h10_q"]
            grid.hits += g["hits"]
            grid.n_batches += g["n_batches"]
            grid.n_histories += g["n_histories"]
    if per_process:
        return out, grid, res
    return out, grid
