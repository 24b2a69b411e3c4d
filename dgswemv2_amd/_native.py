"""ctypes view of include/rkdg.h and the loader of librkdg.so.

The library is built in-tree (dgswemv2_amd/librkdg.so) by __graft_entry__.build(); loading
fails loudly when it is missing -- there is no fallback implementation.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "librkdg.so"

# rkdg_status
RKDG_OK, RKDG_ERR_ARG, RKDG_ERR_HIP, RKDG_ERR_NAN, RKDG_ERR_UNSUPPORTED, RKDG_ERR_NO_DEVICE = range(6)
STATUS_NAMES = {0: "RKDG_OK", 1: "RKDG_ERR_ARG", 2: "RKDG_ERR_HIP", 3: "RKDG_ERR_NAN",
                4: "RKDG_ERR_UNSUPPORTED", 5: "RKDG_ERR_NO_DEVICE"}
# rkdg_precision
RKDG_FP64_PLAIN, RKDG_FP32_BULK, RKDG_FP64_FAST = 0, 1, 2
# boundary codes
BC_LAND, BC_TIDE, BC_FLOW, BC_FUNCTION, BC_LEVEE = 0, 1, 2, 3, 129
# problem function sets
PROBLEM_NONE, PROBLEM_MANUFACTURED, PROBLEM_PARABOLIC_BOWL, PROBLEM_DAM_BREAK = 0, 1, 2, 3
FRICTION_NONE, FRICTION_CHEZY, FRICTION_MANNING = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint8)


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("n_elements", C.c_int32),
        ("node_xyz", _dp), ("elt_nodes", _ip),
        ("n_bound_edges", C.c_int32), ("bound_edge_nodes", _ip), ("bound_edge_type", _up),
        ("n_constituents", C.c_int32),
        ("frequency", _dp), ("forcing_fact", _dp), ("equilib_arg", _dp),
        ("node_amplitude", _dp), ("node_phase", _dp), ("node_manning_n", _dp),
        ("n_levee_pairs", C.c_int32), ("levee_nodes", _ip), ("levee_param", _dp),
    ]


class Params(C.Structure):
    _fields_ = [
        ("p", C.c_int32), ("nstages", C.c_int32), ("order", C.c_int32),
        ("dt", C.c_double), ("g", C.c_double), ("ramp_duration", C.c_double),
        ("problem", C.c_int32), ("ic_constant", C.c_int32),
        ("ze0", C.c_double), ("qx0", C.c_double), ("qy0", C.c_double),
        ("function_source", C.c_int32), ("friction", C.c_int32), ("friction_cf", C.c_double),
        ("wetting_drying", C.c_int32), ("h_o", C.c_double),
        ("slope_limiting", C.c_int32), ("sl_M", C.c_double), ("sl_nu", C.c_double),
        ("problem_param", C.c_double * 4),
        ("spherical_mesh", C.c_int32),
        ("sp_correction", C.c_int32),
        ("sph_R", C.c_double),
        ("sph_lon_o", C.c_double),
        ("sph_lat_o", C.c_double),
    ]


class PerfReport(C.Structure):
    _fields_ = [
        ("n_elements", C.c_int64), ("p", C.c_int32), ("ndof", C.c_int32), ("nstages", C.c_int32),
        ("precision", C.c_int32), ("stages_timed", C.c_int64),
        ("stage_kernel_ms", C.c_double), ("limiter_kernel_ms", C.c_double), ("total_ms", C.c_double),
        ("stage_bytes_alg", C.c_double), ("regs_per_thread", C.c_int32),
    ]


class Schedule(C.Structure):
    _fields_ = [("block", C.c_int32), ("groups_per_xcd", C.c_int32), ("graph", C.c_int32)]


class QuadMeshSpec(C.Structure):
    _fields_ = [
        ("x1", C.c_double), ("y1", C.c_double), ("x2", C.c_double), ("y2", C.c_double),
        ("m", C.c_int32), ("n", C.c_int32), ("pattern", C.c_int32), ("bc", C.c_int32 * 4),
        ("bathy_kind", C.c_int32), ("b0", C.c_double), ("alpha", C.c_double),
        ("bmin", C.c_double), ("bmax", C.c_double), ("n_modes", C.c_int32),
        ("jitter", C.c_double), ("seed", C.c_uint64),
        ("step_x", C.c_double), ("step_b_left", C.c_double), ("step_b_right", C.c_double),
        ("frequency", C.c_double), ("forcing_fact", C.c_double), ("equilib_arg", C.c_double),
        ("amplitude", C.c_double), ("phase", C.c_double), ("manning_n", C.c_double),
    ]


MLP_IN, MLP_HIDDEN = 36, 64
MLP_NPARAM = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + 2 * (MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN) + MLP_HIDDEN + 1


class MlpWeights(C.Structure):
    _fields_ = [
        ("w1", C.c_float * (MLP_HIDDEN * MLP_IN)), ("b1", C.c_float * MLP_HIDDEN),
        ("w2", C.c_float * (MLP_HIDDEN * MLP_HIDDEN)), ("b2", C.c_float * MLP_HIDDEN),
        ("w3", C.c_float * (MLP_HIDDEN * MLP_HIDDEN)), ("b3", C.c_float * MLP_HIDDEN),
        ("w4", C.c_float * MLP_HIDDEN), ("b4", C.c_float), ("threshold", C.c_float),
    ]


class IndicatorStats(C.Structure):
    _fields_ = [
        ("stages", C.c_int64), ("candidates", C.c_int64), ("flagged", C.c_int64), ("write_backs", C.c_int64),
        ("significant", C.c_int64), ("missed", C.c_int64), ("missed_significant", C.c_int64),
        ("max_missed_change", C.c_double),
    ]


class TrainParams(C.Structure):
    _fields_ = [
        ("batch", C.c_int32), ("epochs", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
        ("beta2", C.c_float), ("eps", C.c_float), ("pos_weight", C.c_float), ("seed", C.c_uint64),
    ]


_fp = C.POINTER(C.c_float)

# Every entry point declared in include/rkdg.h, with its ctypes signature.
SIGNATURES = {
    "rkdg_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)]),
    "rkdg_destroy": (None, [C.c_void_p]),
    "rkdg_last_error": (C.c_char_p, [C.c_void_p]),
    "rkdg_create_error": (C.c_char_p, []),
    "rkdg_set_state": (C.c_int, [C.c_void_p, _dp, _up]),
    "rkdg_get_state": (C.c_int, [C.c_void_p, _dp, _up, _up, _up]),
    "rkdg_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "rkdg_stage": (C.c_int, [C.c_void_p]),
    "rkdg_get_time": (C.c_int, [C.c_void_p, _dp, _ip, _ip]),
    "rkdg_set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    "rkdg_nan_info": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "rkdg_sync": (C.c_int, [C.c_void_p]),
    "rkdg_l2_residual": (C.c_int, [C.c_void_p, _dp]),
    "rkdg_report": (C.c_int, [C.c_void_p, C.POINTER(PerfReport)]),
    "rkdg_reset_timers": (C.c_int, [C.c_void_p]),
    "rkdg_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "rkdg_master_tables": (C.c_int, [C.c_int32, _ip, _ip, _dp, _dp, _dp, _dp]),
    "rkdg_ssprk_coeffs": (C.c_int, [C.c_int32, C.c_int32, _dp, _dp, _dp]),
    "rkdg_mesh_generate": (C.c_int, [C.POINTER(QuadMeshSpec), C.POINTER(C.c_void_p), C.POINTER(MeshDesc)]),
    "rkdg_mesh_free": (None, [C.c_void_p]),
    "rkdg_mesh_read_adcirc": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(MeshDesc)]),
    "rkdg_mesh_write_adcirc": (C.c_int, [C.POINTER(MeshDesc), C.c_char_p, C.c_char_p, C.c_char_p]),
    "rkdg_mesh_write_meta": (C.c_int, [C.POINTER(MeshDesc), C.c_char_p, C.c_char_p]),
    "rkdg_partition_rcb": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, C.c_void_p]),
    "rkdg_submesh": (C.c_int, [C.POINTER(MeshDesc), C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(MeshDesc),
                               C.POINTER(C.POINTER(C.c_int32))]),
    "rkdg_mesh_renumber": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p), C.POINTER(MeshDesc),
                                     C.POINTER(C.POINTER(C.c_int32))]),
    "rkdg_stage_begin": (C.c_int, [C.c_void_p]),
    "rkdg_stage_finish": (C.c_int, [C.c_void_p]),
    "rkdg_halo_define": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]),
    "rkdg_halo_buffer": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    "rkdg_halo_pack": (C.c_int, [C.c_void_p, C.c_int32]),
    "rkdg_halo_unpack": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    "rkdg_mesh_read_meta": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(MeshDesc)]),
    "rkdg_create_batch": (C.c_int, [C.c_int32, C.POINTER(MeshDesc), C.POINTER(Params), C.c_int,
                                    C.POINTER(C.c_void_p)]),
    "rkdg_batch_size": (C.c_int32, [C.c_void_p]),
    "rkdg_batch_get_state": (C.c_int, [C.c_void_p, C.c_int32, _dp, _up, _up, _up]),
    "rkdg_batch_set_state": (C.c_int, [C.c_void_p, C.c_int32, _dp, _up]),
    "rkdg_batch_get_time": (C.c_int, [C.c_void_p, C.c_int32, _dp, _ip, _ip]),
    "rkdg_batch_nan_info": (C.c_int, [C.c_void_p, _ip, _ip, _ip, _ip]),
    "rkdg_batch_set_schedule": (C.c_int, [C.c_void_p, C.POINTER(Schedule)]),
    "rkdg_batch_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(Schedule)]),
    "rkdg_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Schedule), _dp]),
    "rkdg_write_modal": (C.c_int, [C.c_void_p, C.c_char_p]),
    "rkdg_write_vtu": (C.c_int, [C.c_void_p, C.c_char_p]),
    "rkdg_write_vtk": (C.c_int, [C.c_void_p, C.c_char_p]),
    "rkdg_set_indicator": (C.c_int, [C.c_void_p, C.POINTER(MlpWeights), C.c_int32]),
    "rkdg_get_indicator_stats": (C.c_int, [C.c_void_p, C.POINTER(IndicatorStats), C.c_int32]),
    "rkdg_capture": (C.c_int, [C.c_void_p, C.c_int32]),
    "rkdg_get_capture": (C.c_int, [C.c_void_p, _fp, _up, _up, _fp, _fp]),
    "rkdg_mlp_trainer_create": (C.c_int, [C.POINTER(MlpWeights), C.c_int32, C.POINTER(C.c_void_p)]),
    "rkdg_mlp_set_data": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32]),
    "rkdg_mlp_train": (C.c_int, [C.c_void_p, C.POINTER(TrainParams), _fp]),
    "rkdg_mlp_get_weights": (C.c_int, [C.c_void_p, C.POINTER(MlpWeights)]),
    "rkdg_mlp_forward": (C.c_int, [C.c_void_p, _fp, C.c_int32, _fp]),
    "rkdg_mlp_forward_noalloc": (C.c_int, [C.POINTER(MlpWeights), _fp, C.c_int32, _fp]),
    "rkdg_mlp_gradient": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, C.c_float, _fp, C.POINTER(MlpWeights)]),
    "rkdg_mlp_trainer_error": (C.c_char_p, [C.c_void_p]),
    "rkdg_mlp_trainer_destroy": (None, [C.c_void_p]),
}

_lib = None


def load_library(path=None):
    """Load librkdg.so (built in-tree).  Raises OSError when it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise OSError(f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def uptr(a):
    return a.ctypes.data_as(_up) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def weights_from_flat(flat, threshold=0.0):
    """MlpWeights from the flat parameter vector (w1 b1 w2 b2 w3 b3 w4 b4)."""
    flat = np.ascontiguousarray(flat, np.float32).reshape(-1)
    n = C.sizeof(MlpWeights) // 4 - 1
    if flat.size != n:
        raise ValueError(f"expected {n} parameters, got {flat.size}")
    w = MlpWeights()
    C.memmove(C.addressof(w), flat.ctypes.data, 4 * n)
    w.threshold = threshold
    return w


def weights_to_flat(w):
    n = C.sizeof(MlpWeights) // 4 - 1
    return np.frombuffer(bytes(w), np.float32, count=n).copy()


class RkdgError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
