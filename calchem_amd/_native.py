"""ctypes binding of libcalchem_hip.so (the C ABI in include/calchem_hip.h).

The library is the only compute path: if it is missing, or no GPU is
visible, every entry point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcalchem_hip.so")

STATUS = {
    0: "ok", -1: "bad arguments", -2: "shape or handle mismatch", -3: "no usable device",
    -4: "allocation failure", -5: "launch or copy failure", -6: "size limit exceeded",
    -7: "unsupported option",
}


class CalchemHipError(RuntimeError):
    def __init__(self, what: str, code: int):
        super().__init__(f"{what} failed with {code} ({STATUS.get(code, 'unknown')})")
        self.code = code


class VxcOptions(C.Structure):
    _fields_ = [("block_points", C.c_int32), ("group_points", C.c_int32),
                ("memory_limit_bytes", C.c_size_t), ("mixed_exact_threshold", C.c_double),
                ("rho_tile_points", C.c_int32), ("group_max_rows", C.c_int32), ("ao_cache", C.c_int32),
                ("v_small_bf16", C.c_int32)]


class VxcStats(C.Structure):
    _fields_ = [("n_points_in", C.c_int64), ("n_points", C.c_int64), ("n_blocks", C.c_int64),
                ("n_groups", C.c_int64), ("n_rho_tiles", C.c_int64), ("n_v_pairs", C.c_int64),
                ("flops_rho", C.c_double), ("flops_v", C.c_double),
                ("flops_rho_executed_fp64", C.c_double), ("flops_rho_executed_fp32", C.c_double),
                ("flops_v_executed", C.c_double),
                ("max_m_rho", C.c_int32), ("max_m_v", C.c_int32), ("rho_points_fp64", C.c_int32),
                ("rho_points_fp32", C.c_int32), ("device_bytes", C.c_size_t),
                ("lds_bytes_per_block", C.c_size_t),
                ("profiled_builds", C.c_int64), ("ms_gather", C.c_double), ("ms_rho", C.c_double),
                ("ms_xc", C.c_double), ("ms_vxc", C.c_double), ("ms_reduce_v", C.c_double), ("ms_reduce_en", C.c_double),
                ("flops_rho_fp64_rows", C.c_double), ("ao_cache_bytes", C.c_size_t)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTS = [
    "calchem_hip_is_available", "calchem_hip_get_memory_info", "calchem_hip_total_allocated_bytes",
    "calchem_hip_vxc_upload_grid", "calchem_hip_vxc_eval_rho", "calchem_hip_vxc_accum",
    "calchem_hip_vxc_release", "calchem_hip_vxc_create", "calchem_hip_vxc_build",
    "calchem_hip_vxc_build_device", "calchem_hip_vxc_build_uks", "calchem_hip_vxc_build_uks_device",
    "calchem_hip_vxc_create_batched", "calchem_hip_vxc_build_batched", "calchem_hip_vxc_build_batched_device",
    "calchem_hip_vxc_graph_create", "calchem_hip_vxc_graph_launch", "calchem_hip_vxc_graph_destroy",
    "calchem_hip_vxc_set_atoms", "calchem_hip_vxc_xc_gradient",
    "calchem_hip_vxc_set_profiling", "calchem_hip_vxc_stats",
    "calchem_hip_grid_build", "calchem_hip_grid_streaming_tasks", "calchem_hip_grid_task_points",
    "calchem_hip_xc_eval", "calchem_hip_xc_eval_device",
    "calchem_hip_xc_eval_device_noalloc",
    "calchem_hip_int_create", "calchem_hip_int_release", "calchem_hip_int_one_electron", "calchem_hip_int_eri_diag",
    "calchem_hip_int_eri_columns", "calchem_hip_int_cholesky", "calchem_hip_int_jk",
    "calchem_hip_mlp_create", "calchem_hip_mlp_release", "calchem_hip_mlp_get_params", "calchem_hip_mlp_forward",
    "calchem_hip_mlp_loss_grad", "calchem_hip_mlp_set_data", "calchem_hip_mlp_train",
]

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (building nothing; run build_native.build() first)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CalchemHipError(f"libcalchem_hip.so not found at {LIB_PATH}; build it first", -3)
    L = C.CDLL(LIB_PATH)
    vp, i32, u64, dp = C.c_void_p, C.c_int, C.c_uint64, C.c_void_p
    L.calchem_hip_is_available.restype = C.c_int
    L.calchem_hip_get_memory_info.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.calchem_hip_total_allocated_bytes.restype = C.c_size_t
    L.calchem_hip_vxc_upload_grid.argtypes = [vp, i32, vp, i32, vp, i32, vp, i32, vp, i32, vp, i32, i32,
                                              u64, C.c_size_t]
    L.calchem_hip_vxc_eval_rho.argtypes = [u64, dp, dp, i32, i32]
    L.calchem_hip_vxc_accum.argtypes = [u64, dp, dp, i32, i32]
    L.calchem_hip_vxc_release.argtypes = [u64]
    L.calchem_hip_vxc_create.argtypes = [vp, i32, vp, i32, vp, i32, dp, i32, i32,
                                         C.POINTER(VxcOptions), u64]
    L.calchem_hip_vxc_build.argtypes = [u64, i32, i32, dp, dp, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
    L.calchem_hip_vxc_build_device.argtypes = [u64, i32, i32, dp, dp, dp, vp]
    L.calchem_hip_vxc_build_uks.argtypes = [u64, i32, i32, dp, dp, dp, dp, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]
    L.calchem_hip_vxc_build_uks_device.argtypes = [u64, i32, i32, dp, dp, dp, dp, dp, vp]
    L.calchem_hip_vxc_create_batched.argtypes = [i32, vp, vp, vp, vp, vp, i32, vp, i32, dp, vp, u64]
    L.calchem_hip_vxc_build_batched.argtypes = [u64, i32, i32, dp, dp, dp, dp]
    L.calchem_hip_vxc_build_batched_device.argtypes = [u64, i32, i32, dp, dp, dp, vp]
    L.calchem_hip_vxc_graph_create.argtypes = [u64, i32, i32, dp, dp, dp, C.POINTER(C.c_uint64)]
    L.calchem_hip_vxc_graph_launch.argtypes = [u64, u64, vp]
    L.calchem_hip_vxc_graph_destroy.argtypes = [u64, u64]
    L.calchem_hip_vxc_set_atoms.argtypes = [u64, i32, vp, vp, vp, vp]
    L.calchem_hip_vxc_xc_gradient.argtypes = [u64, i32, vp, vp, vp]
    L.calchem_hip_vxc_set_profiling.argtypes = [u64, i32]
    L.calchem_hip_vxc_stats.argtypes = [u64, C.POINTER(VxcStats)]
    L.calchem_hip_grid_build.restype = C.c_long
    L.calchem_hip_grid_build.argtypes = [i32, vp, vp, i32, i32, i32, i32, vp, C.c_long]
    L.calchem_hip_grid_streaming_tasks.restype = C.c_long
    L.calchem_hip_grid_streaming_tasks.argtypes = [i32, vp, i32, i32, i32, vp, C.c_long]
    L.calchem_hip_grid_task_points.restype = C.c_long
    L.calchem_hip_grid_task_points.argtypes = [i32, vp, vp, i32, i32, i32, i32, i32, i32, i32, vp, C.c_long]
    L.calchem_hip_xc_eval.argtypes = [i32, C.c_longlong, vp, vp, vp, vp]
    L.calchem_hip_xc_eval_device.argtypes = [i32, C.c_longlong, vp, vp, vp, vp, vp]
    L.calchem_hip_xc_eval_device_noalloc.argtypes = [i32, C.c_longlong, vp, vp, vp, vp, vp]
    L.calchem_hip_int_create.argtypes = [vp, i32, vp, i32, vp, i32, i32, i32, vp, u64]
    L.calchem_hip_int_release.argtypes = [u64]
    L.calchem_hip_int_one_electron.argtypes = [u64, vp, vp, vp]
    L.calchem_hip_int_eri_diag.argtypes = [u64, vp]
    L.calchem_hip_int_eri_columns.argtypes = [u64, i32, vp, vp, vp]
    L.calchem_hip_int_cholesky.argtypes = [u64, C.c_double, i32, C.POINTER(C.c_int)]
    L.calchem_hip_int_jk.argtypes = [u64, i32, vp, i32, vp, vp, vp]
    L.calchem_hip_mlp_create.argtypes = [i32, i32, i32, vp, u64]
    L.calchem_hip_mlp_release.argtypes = [u64]
    L.calchem_hip_mlp_get_params.argtypes = [u64, vp]
    L.calchem_hip_mlp_forward.argtypes = [u64, i32, vp, vp]
    L.calchem_hip_mlp_loss_grad.argtypes = [u64, i32, vp, vp, vp, vp, vp]
    L.calchem_hip_mlp_set_data.argtypes = [u64, i32, vp, vp, vp]
    L.calchem_hip_mlp_train.argtypes = [u64, i32, i32, C.c_float, u64, vp]
    _lib = L
    return L


def check(code: int, what: str) -> int:
    if code < 0:
        raise CalchemHipError(what, int(code))
    return code


def is_available() -> bool:
    return load().calchem_hip_is_available() == 1


def require_device() -> C.CDLL:
    L = load()
    if L.calchem_hip_is_available() != 1:
        raise CalchemHipError("calchem_hip_is_available (no MI355X visible, or CALCHEM_HIP_DISABLE=1)", -3)
    return L


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


_next_handle = [0x5EED0000]


def new_handle() -> int:
    _next_handle[0] += 1
    return _next_handle[0]
