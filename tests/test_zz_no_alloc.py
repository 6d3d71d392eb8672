"""cc_transport_noalloc (the allocation-free variant of cc_transport) against cc_transport.

Both are called through the C-ABI on the same photons, seed and bound grid: the variant keeps its
photon records, queues and staging in the library's static per-GPU scratch array instead of the
context's pool, and runs the same kernels, so every per-photon output and counter is equal and
the FP64 sums (deposits, grid) agree to summation order.  Two calls in a row, the first at the
variant's maximum size and the second smaller, check that a call does not depend on what the
previous one left in the shared array.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene():
    from chatcarlo_amd.geometry import Geometry
    return Geometry([{"name": "water", "shape": "box", "material": "water", "center": [0.0, 0.0, 0.0],
                      "size_cm": [20.0, 30.0, 30.0]},
                     {"name": "lead", "shape": "box", "material": "lead", "center": [12.0, 0.0, 0.0],
                      "size_cm": [0.2, 30.0, 30.0]},
                     {"name": "bone", "shape": "cylinder", "material": "bone", "center": [0.0, 0.0, 0.0],
                      "radius_cm": 2.0, "height_cm": 20.0, "axis": "z"}])


def _photons(n, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 3))
    pos[:, 0] = -15.0
    pos[:, 1:] = rng.uniform(-5.0, 5.0, (n, 2))
    d = np.zeros((n, 3))
    d[:, 0] = 1.0
    d[:, 1:] = rng.normal(0.0, 0.1, (n, 2))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    e = rng.uniform(20.0, 120.0, n)
    return pos, d, e


def _call(eng, fn, pos, d, e, seed, first, flags):
    from chatcarlo_amd import _native as N
    n = len(e)
    p, dd, ee = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (pos, d, e))
    nsc = np.zeros(n, np.int32)
    ab = np.zeros(n, np.uint8)
    es = np.zeros(n, np.uint8)
    out = N.CcRunOut()
    rc = fn(eng.ctx, n, N.ptr(p, C.c_double), N.ptr(dd, C.c_double), N.ptr(ee, C.c_double), seed, first, flags,
            N.ptr(nsc, C.c_int32), N.ptr(ab, C.c_uint8), N.ptr(es, C.c_uint8), C.byref(out))
    N.check(rc)
    return p, dd, ee, nsc, ab, es, out


COUNTERS = ("n_histories", "n_absorbed", "n_escaped", "scatter_sum", "n_fluorescence", "n_events", "n_flights",
            "n_voxel_visits", "error_bits")


def test_noalloc_transport_equals_cc_transport(engine):
    from chatcarlo_amd import _native as N
    lib = engine.lib
    engine.set_geometry(_scene())
    origin, shape, vox = (-15.0, -15.0, -15.0), (30, 30, 30), 1.0
    flags = N.RUN_FLUORESCENCE | N.RUN_GRID
    for n, seed, first in ((N.NOALLOC_MAX_PHOTONS, 11, 0), (3000, 12, 777)):
        pos, d, e = _photons(n, seed)
        print(f"noalloc vs cc_transport: n={n} seed={seed} first={first}", flush=True)
        res = []
        for fn in (lib.cc_transport, lib.cc_transport_noalloc):
            engine.bind_grid(origin, shape, vox, False)
            r = _call(engine, fn, pos, d, e, seed, first, flags)
            g = engine.read_grid(shape, moments=False)
            res.append((r, g))
        (a, ga), (b, gb) = res
        for x, y in zip(a[:6], b[:6]):
            assert np.array_equal(x, y)
        for k in COUNTERS:
            assert getattr(a[6], k) == getattr(b[6], k), k
        assert a[6].n_histories == n and a[6].n_fluorescence > 0 and a[6].n_voxel_visits > 0
        ea, eb = np.array(a[6].edep_keV), np.array(b[6].edep_keV)
        assert np.allclose(ea, eb, rtol=1e-12, atol=0.0)
        for key in ("kerma", "h10"):
            assert ga[key].max() > 0.0
            assert np.allclose(ga[key], gb[key], rtol=1e-12, atol=1e-12 * float(ga[key].max()))


def test_noalloc_transport_rejects_what_it_cannot_hold(engine):
    from chatcarlo_amd import _native as N
    lib = engine.lib
    engine.set_geometry(_scene())
    pos, d, e = _photons(N.NOALLOC_MAX_PHOTONS + 1, 3)
    with pytest.raises(ValueError, match="at most"):
        _call(engine, lib.cc_transport_noalloc, pos, d, e, 1, 0, N.RUN_FLUORESCENCE)
    pos, d, e = _photons(16, 4)
    with pytest.raises(ValueError, match="CC_RUN_TRACE"):
        _call(engine, lib.cc_transport_noalloc, pos, d, e, 1, 0, N.RUN_FLUORESCENCE | N.RUN_TRACE)
