"""ctypes binding of libchatcarlo_hip.so (include/chatcarlo_hip.h).

The library is built in-tree by chatcarlo_amd.build; importing this module
never falls back to anything else: a missing library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libchatcarlo_hip.so"

CC_OK, CC_ERR_ARG, CC_ERR_HIP, CC_ERR_STATE, CC_ERR_DEVICE = 0, -1, -2, -3, -4
CC_MAX_PRIMS, CC_MAX_MAT = 64, 32
RUN_FLUORESCENCE, RUN_GRID, RUN_TRACK_UNCERTAINTY, RUN_TIMING = 1, 2, 4, 8
RUN_REGION, RUN_RECORD, RUN_IMPORTANCE, RUN_SOURCE_BIAS = 16, 32, 64, 128
RUN_TRACE, RUN_PER_BATCH, RUN_PERSISTENT = 256, 512, 1024
NOALLOC_MAX_PHOTONS = 1 << 16   # photons per cc_transport_noalloc call
EV_BOUNDARY, EV_ESCAPE, EV_PHOTOELECTRIC, EV_FLUORESCENCE, EV_COMPTON, EV_RAYLEIGH = range(6)
TRACE_DTYPE = np.dtype([("photon", np.uint32), ("flight", np.int32), ("step", np.int32), ("material", np.int32),
                        ("start", np.float32, 3), ("end", np.float32, 3), ("energy_keV", np.float32),
                        ("event", np.int32)])
MLP_IN, MLP_HIDDEN = 16, 64
MLP_PARAMS = MLP_IN * MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1
RECORD_FLOATS = 12
SRC_CELLS = 512
MAP_FLIGHT, MAP_COLLISION = 0, 1
DEV_ENERGY_RANGE, DEV_ITER_CAP, DEV_Q_RANGE, DEV_STEP_CAP, DEV_ROUND_CAP = 1, 2, 4, 8, 16

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
u64p = C.POINTER(C.c_uint64)


class CcPrim(C.Structure):
    _fields_ = [("kind", C.c_int32), ("axis", C.c_int32), ("material", C.c_int32), ("pad", C.c_int32),
                ("p", C.c_float * 8)]


class CcWorld(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("hi", C.c_float * 3), ("background", C.c_int32),
                ("n_prim", C.c_int32)]


class CcTables(C.Structure):
    _fields_ = [("n_mat", C.c_int32), ("n_e", C.c_int32), ("n_el", C.c_int32), ("n_q", C.c_int32),
                ("max_edges", C.c_int32), ("n_h10", C.c_int32), ("sel_len", C.c_int32),
                ("elem_len", C.c_int32),
                ("density", fp), ("e_node", fp), ("dl_node", fp), ("sig", fp), ("slope", fp),
                ("edge_of_interval", ip), ("edge_e", fp), ("edge_dl", fp), ("edge_sig", fp),
                ("edge_slope", fp), ("mat_elem_off", ip), ("mat_elem", ip), ("sel_off", ip),
                ("sel", fp), ("sq_over_z", fp), ("q_s", fp), ("ray_x", fp), ("ray_a", fp),
                ("fl", fp), ("h10", fp)]


class CcSource(C.Structure):
    _fields_ = [("field", C.c_int32), ("n_spec", C.c_int32), ("n_angles", C.c_int32),
                ("rotation", C.c_int32),
                ("pos", C.c_double * 3), ("dir", C.c_double * 3), ("u", C.c_double * 3),
                ("v", C.c_double * 3), ("sid", C.c_double), ("width", C.c_double),
                ("height", C.c_double), ("cos_half", C.c_double), ("iso", C.c_double * 3),
                ("axis", C.c_double * 3), ("scan_length", C.c_double), ("spec_e", dp),
                ("spec_cdf", dp), ("heel_bins", C.c_int32), ("heel_nz_lo", C.c_int32),
                ("heel_nz_hi", C.c_int32), ("heel_pad", C.c_int32), ("heel_hu", C.c_double),
                ("heel_hv", C.c_double), ("heel_span", C.c_double), ("heel_rel_flu", dp), ("heel_cdf", dp)]


class CcRunOut(C.Structure):
    _fields_ = [("n_histories", C.c_uint64), ("n_absorbed", C.c_uint64), ("n_escaped", C.c_uint64),
                ("scatter_sum", C.c_uint64), ("n_fluorescence", C.c_uint64),
                ("n_batches", C.c_uint64), ("n_events", C.c_uint64), ("n_flights", C.c_uint64),
                ("n_voxel_visits", C.c_uint64), ("error_bits", C.c_uint32),
                ("rounds_max", C.c_uint32), ("edep_keV", C.c_double * CC_MAX_MAT),
                ("edep_sum2", C.c_double * CC_MAX_MAT), ("kernel_ms", C.c_double * 8),
                ("kernel_launches", C.c_uint64 * 8),
                ("region_sum", C.c_double), ("region_sum2", C.c_double),
                ("region_batch_sum2", C.c_double), ("n_split", C.c_uint64),
                ("n_roulette", C.c_uint64), ("n_records", C.c_uint64),
                ("kernel_busy_ms", C.c_double * 8)]


class CcTrainCfg(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("yscale", C.c_float), ("lr_end", C.c_float),
                ("seed", C.c_uint64), ("target_power", C.c_uint32), ("pad", C.c_uint32)]


class CcTrainOut(C.Structure):
    _fields_ = [("loss_first", C.c_double), ("loss_last", C.c_double), ("ms", C.c_double),
                ("yscale", C.c_float), ("n_records", C.c_uint32), ("n_positive", C.c_uint32),
                ("pad", C.c_uint32)]


class CcImportanceCfg(C.Structure):
    _fields_ = [("cw", C.c_float), ("w_lo", C.c_float), ("w_hi", C.c_float), ("split_max", C.c_int32),
                ("pool_factor", C.c_float), ("source_bias", C.c_int32), ("bias_mix", C.c_float),
                ("samples_per_cell", C.c_int32), ("use_post", C.c_int32), ("use_pre", C.c_int32),
                ("split_gain", C.c_float), ("pad", C.c_uint32), ("seed", C.c_uint64)]


class CcTraceRecord(C.Structure):
    _fields_ = [("photon", C.c_uint32), ("flight", C.c_int32), ("step", C.c_int32), ("material", C.c_int32),
                ("start", C.c_float * 3), ("end", C.c_float * 3), ("energy_keV", C.c_float), ("event", C.c_int32)]


class CcBatchTuning(C.Structure):
    _fields_ = [("block", C.c_int32), ("rounds_per_sync", C.c_int32), ("pool", C.c_uint32),
                ("tally_copies", C.c_uint32), ("persistent", C.c_int32), ("epoch", C.c_int32)]


class CcBatchOut(C.Structure):
    _fields_ = [("n_histories", C.c_uint64), ("n_absorbed", C.c_uint64), ("n_escaped", C.c_uint64),
                ("scatter_sum", C.c_uint64), ("n_fluorescence", C.c_uint64), ("n_events", C.c_uint64),
                ("n_flights", C.c_uint64), ("n_voxel_visits", C.c_uint64), ("rounds", C.c_uint64),
                ("graph_launches", C.c_uint64), ("device_ms", C.c_double), ("error_bits", C.c_uint32),
                ("pad", C.c_uint32)]


_SIGS = {
    "cc_last_error": (C.c_char_p, []),
    "cc_version": (C.c_int, [ip, ip]),
    "cc_ctx_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "cc_ctx_destroy": (C.c_int, [C.c_void_p]),
    "cc_upload_tables": (C.c_int, [C.c_void_p, C.POINTER(CcTables)]),
    "cc_upload_scene": (C.c_int, [C.c_void_p, C.POINTER(CcPrim), C.c_int32, C.POINTER(CcWorld)]),
    "cc_bind_grid": (C.c_int, [C.c_void_p, dp, ip, C.c_double, C.c_int32]),
    "cc_unbind_grid": (C.c_int, [C.c_void_p]),
    "cc_set_grid_copies": (C.c_int, [C.c_void_p, C.c_int32]),
    "cc_atomic_rate": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "cc_trace_alloc": (C.c_int, [C.c_void_p, C.c_uint32]),
    "cc_trace_read": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, u64p]),
    "cc_reset_grid": (C.c_int, [C.c_void_p]),
    "cc_read_grid": (C.c_int, [C.c_void_p, dp, dp, dp, dp, ip, i64p, i64p]),
    "cc_load_grid": (C.c_int, [C.c_void_p, dp, dp, dp, dp, ip, C.c_int64, C.c_int64]),
    "cc_run": (C.c_int, [C.c_void_p, C.POINTER(CcSource), C.c_uint64, C.c_uint64, C.c_uint64,
                         C.c_uint32, C.c_uint32, C.POINTER(CcRunOut)]),
    "cc_transport": (C.c_int, [C.c_void_p, C.c_int64, dp, dp, dp, C.c_uint64, C.c_uint64, C.c_uint32,
                               ip, u8p, u8p, C.POINTER(CcRunOut)]),
    "cc_transport_noalloc": (C.c_int, [C.c_void_p, C.c_int64, dp, dp, dp, C.c_uint64, C.c_uint64, C.c_uint32,
                                       ip, u8p, u8p, C.POINTER(CcRunOut)]),
    "cc_locate": (C.c_int, [C.c_void_p, C.c_int64, fp, ip]),
    "cc_test_boundary": (C.c_int, [C.c_void_p, C.c_int64, fp, fp, C.c_float, fp, ip]),
    "cc_test_xs": (C.c_int, [C.c_void_p, C.c_int64, ip, fp, fp]),
    "cc_test_compton": (C.c_int, [C.c_void_p, C.c_int64, ip, fp, fp, C.c_int32, fp, fp, ip, ip]),
    "cc_test_rayleigh": (C.c_int, [C.c_void_p, C.c_int64, ip, fp, fp, C.c_int32, fp, ip, ip]),
    "cc_test_fluor": (C.c_int, [C.c_void_p, C.c_int64, ip, fp, fp, ip, ip, fp]),
    "cc_test_rotate": (C.c_int, [C.c_int64, fp, fp, fp, fp]),
    "cc_test_isotropic": (C.c_int, [C.c_int64, fp, fp, fp]),
    "cc_test_dda": (C.c_int, [C.c_void_p, C.c_int64, fp, fp, fp, dp, dp]),
    "cc_test_philox": (C.c_int, [C.c_int64, u32p, u32p, u32p]),
    "cc_test_source": (C.c_int, [C.c_void_p, C.POINTER(CcSource), C.c_int64, C.c_uint64, C.c_uint64,
                                 fp, fp, fp]),
    "cc_set_region": (C.c_int, [C.c_void_p, dp, C.c_double, C.c_double]),
    "cc_records_alloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_float]),
    "cc_records_read": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, fp, u32p]),
    "cc_records_clear": (C.c_int, [C.c_void_p]),
    "cc_mlp_set_params": (C.c_int, [C.c_void_p, C.c_int32, fp]),
    "cc_mlp_get_params": (C.c_int, [C.c_void_p, C.c_int32, fp]),
    "cc_mlp_features": (C.c_int, [C.c_void_p, C.c_int64, fp, fp, fp, fp]),
    "cc_mlp_forward": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, fp, fp]),
    "cc_mlp_grad": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, fp, fp, C.c_float, fp, dp]),
    "cc_mlp_train": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(CcTrainCfg), C.POINTER(CcTrainOut), dp]),
    "cc_set_importance": (C.c_int, [C.c_void_p, C.POINTER(CcImportanceCfg), C.POINTER(CcSource), fp, fp]),
    "cc_batch_upload": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(CcPrim), ip, C.POINTER(CcWorld),
                                  C.POINTER(CcSource), u64p, u64p, dp, ip, dp]),
    "cc_batch_run": (C.c_int, [C.c_void_p, C.POINTER(CcBatchTuning), C.c_double, C.c_uint32,
                               C.POINTER(CcBatchOut)]),
    "cc_batch_read": (C.c_int, [C.c_void_p, C.c_int32, dp, dp, dp]),
    "cc_batch_autotune": (C.c_int, [C.c_void_p, C.c_double, C.c_uint32, C.POINTER(CcBatchTuning), dp,
                                    C.c_int32, ip]),
}
EXPORTS = tuple(_SIGS)

_lib = None


class NativeError(RuntimeError):
    pass


def load(path: Path | None = None):
    """Load the library (raises if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise NativeError(f"{p} is not built; run chatcarlo_amd.build.build_library() first")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int):
    if rc == CC_OK:
        return
    msg = load().cc_last_error().decode(errors="replace")
    if rc == CC_ERR_ARG:
        raise ValueError(msg)
    raise NativeError(f"chatcarlo_hip error {rc}: {msg}")


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
