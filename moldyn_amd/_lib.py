"""ctypes binding of libmoldyn_hip.so (the C ABI declared in include/moldyn_hip.h).

The library is built in-tree by moldyn_amd.build; if it is missing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmoldyn_hip.so")

MD_OK = 0
MD_ERR_ARG = -1
MD_ERR_HIP = -2
MD_ERR_NOMEM = -3
MD_ERR_UNIMPLEMENTED = -4
MD_ERR_BAROSTAT = -5
MD_ERR_OVERFLOW = -6
MD_ERR_NO_DEVICE = -7

MD_FORCE_FP64_ALLPAIRS = 0
MD_FORCE_FP64_LIST = 1
MD_FORCE_FP32_LIST = 2
MD_FORCE_MLP = 3
MD_MLP_NUM_PARAMS = 20865

# every entry point of include/moldyn_hip.h
EXPORTS = (
    "md_lj_potential_and_force", "md_new_lennard_jones", "md_create", "md_create_replicas",
    "md_destroy", "md_last_error", "md_global_error", "md_get_options", "md_set_options",
    "md_set_pair_potential", "md_set_pair_potential_lj", "md_upload_state", "md_upload_forces", "md_download_state",
    "md_update_force", "md_apply_boundary_conditions", "md_step", "md_step_replicas",
    "md_thermostat_calculate_lambda", "md_thermostat_update", "md_barostat_calculate_myu",
    "md_barostat_update", "md_macro", "md_init_lattice", "md_init_mb_velocities", "md_get_stats",
    "md_set_profiling", "md_kernel_time", "md_mlp_set_params", "md_mlp_get_stats",
    "md_trainer_create", "md_trainer_destroy", "md_trainer_last_error", "md_trainer_feature_stats",
    "md_trainer_set_params", "md_trainer_get_params", "md_trainer_loss", "md_trainer_epoch",
    "md_trainer_predict", "md_frame_writer_create", "md_frame_writer_save", "md_frame_writer_flush",
    "md_frame_writer_stats", "md_frame_writer_last_error", "md_frame_writer_destroy", "md_frame_load",
    "md_text_create", "md_text_destroy", "md_text_last_error", "md_text_table", "md_csv_write_table",
    "md_init_positions", "md_init_velocities", "md_macro_noalloc",
)


class MdThermostat(C.Structure):
    _fields_ = [("kind", C.c_int), ("tau", C.c_double), ("target_t", C.c_double),
                ("lambda_", C.c_double), ("psi", C.c_double)]


class MdBarostat(C.Structure):
    _fields_ = [("kind", C.c_int), ("beta", C.c_double), ("tau", C.c_double),
                ("target_p", C.c_double), ("myu", C.c_double)]


class MdOptions(C.Structure):
    _fields_ = [("force_mode", C.c_int), ("skin", C.c_double), ("block_size", C.c_int),
                ("list_capacity", C.c_int), ("cell_capacity", C.c_int),
                ("cells_per_atom", C.c_double), ("graph_chunk", C.c_int),
                ("tiles_per_block", C.c_int)]


class MdStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("launches", C.c_int64), ("rebuilds", C.c_int64),
                ("refresh_steps", C.c_int64), ("retries", C.c_int64),
                ("list_capacity", C.c_int32), ("cell_capacity", C.c_int32),
                ("mean_list_len", C.c_double), ("max_list_len", C.c_int32),
                ("n_slots", C.c_int32), ("grows", C.c_int64), ("list_visits", C.c_int64),
                ("split_steps", C.c_int64)]


class MdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"moldyn_hip error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load the HIP library; raise ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m moldyn_amd.build`")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    i64p = C.POINTER(C.c_int64)
    vp = C.c_void_p
    L.md_lj_potential_and_force.argtypes = [C.c_double] * 5 + [dp, dp]
    L.md_lj_potential_and_force.restype = None
    L.md_new_lennard_jones.argtypes = [C.c_double, C.c_double, dp, dp]
    L.md_new_lennard_jones.restype = None
    L.md_create.argtypes = [C.POINTER(vp), C.c_int, i64p, dp, dp]
    L.md_create_replicas.argtypes = [C.POINTER(vp), C.c_int, C.c_int, i64p, dp, dp]
    L.md_destroy.argtypes = [vp]
    L.md_destroy.restype = None
    L.md_last_error.argtypes = [vp]
    L.md_last_error.restype = C.c_char_p
    L.md_global_error.argtypes = []
    L.md_global_error.restype = C.c_char_p
    L.md_get_options.argtypes = [vp, C.POINTER(MdOptions)]
    L.md_set_options.argtypes = [vp, C.POINTER(MdOptions)]
    L.md_set_pair_potential.argtypes = [vp, C.c_uint16, C.c_uint16, C.c_double, C.c_double]
    L.md_set_pair_potential_lj.argtypes = [vp, C.c_uint16, C.c_uint16] + [C.c_double] * 4
    L.md_upload_state.argtypes = [vp, dp, dp, dp]
    L.md_upload_forces.argtypes = [vp, dp, dp, dp]
    L.md_download_state.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.md_update_force.argtypes = [vp]
    L.md_apply_boundary_conditions.argtypes = [vp]
    tp = C.POINTER(MdThermostat)
    bp = C.POINTER(MdBarostat)
    L.md_step.argtypes = [vp, C.c_double, tp, bp, C.c_int64]
    L.md_step_replicas.argtypes = [vp, C.c_double, tp, bp, C.c_int64]
    L.md_thermostat_calculate_lambda.argtypes = [vp, tp, C.c_double, C.c_uint16]
    L.md_thermostat_update.argtypes = [vp, tp, C.c_double, C.c_uint16]
    L.md_barostat_calculate_myu.argtypes = [vp, bp, C.c_double, C.c_uint16]
    L.md_barostat_update.argtypes = [vp, bp, C.c_uint16]
    L.md_macro.argtypes = [vp, C.c_int, C.c_uint16, dp]
    L.md_macro_noalloc.argtypes = [dp, dp, dp, C.c_int64, C.c_double, dp, dp, vp]
    L.md_init_lattice.argtypes = [C.c_int, dp, i64p, C.c_double, dp]
    L.md_init_mb_velocities.argtypes = [dp, C.c_int64, C.c_double, C.c_double, C.c_uint64]
    L.md_get_stats.argtypes = [vp, C.POINTER(MdStats)]
    L.md_set_profiling.argtypes = [vp, C.c_int]
    L.md_kernel_time.argtypes = [vp, C.c_int, dp, i64p, dp, i64p]
    fp = C.POINTER(C.c_float)
    L.md_mlp_set_params.argtypes = [vp, fp, fp, fp, fp, fp, C.c_double, C.c_double, C.c_int]
    L.md_mlp_get_stats.argtypes = [vp, C.c_int, dp]
    L.md_trainer_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, dp, dp, dp, dp, C.c_double]
    L.md_trainer_destroy.argtypes = [vp]
    L.md_trainer_destroy.restype = None
    L.md_trainer_last_error.argtypes = [vp]
    L.md_trainer_last_error.restype = C.c_char_p
    L.md_trainer_feature_stats.argtypes = [vp, dp]
    L.md_trainer_set_params.argtypes = [vp, fp, fp, fp, C.c_double, C.c_double]
    L.md_trainer_get_params.argtypes = [vp, fp]
    L.md_trainer_loss.argtypes = [vp, C.c_double, C.c_double, C.c_int, dp, fp]
    L.md_trainer_epoch.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32), dp]
    L.md_trainer_predict.argtypes = [vp, C.c_int, dp, dp]
    u64p = C.POINTER(C.c_uint64)
    L.md_init_positions.argtypes = [vp, C.c_int, C.c_uint16, C.c_int, dp, i64p, C.c_double]
    L.md_init_velocities.argtypes = [vp, C.c_int, C.c_uint16, C.c_double, C.c_uint64]
    L.md_frame_writer_create.argtypes = [C.POINTER(vp), vp, C.c_char_p, C.c_int]
    L.md_frame_writer_save.argtypes = [vp, C.c_int64]
    L.md_frame_writer_flush.argtypes = [vp]
    L.md_frame_writer_stats.argtypes = [vp, dp]
    L.md_frame_writer_last_error.argtypes = [vp]
    L.md_frame_writer_last_error.restype = C.c_char_p
    L.md_frame_writer_destroy.argtypes = [vp]
    L.md_frame_writer_destroy.restype = None
    L.md_frame_load.argtypes = [C.c_char_p, C.c_int64, C.c_int64, i64p, C.POINTER(C.c_uint16), dp, dp, dp]
    L.md_text_create.argtypes = [C.POINTER(vp)]
    L.md_text_destroy.argtypes = [vp]
    L.md_text_destroy.restype = None
    L.md_text_last_error.argtypes = [vp]
    L.md_text_last_error.restype = C.c_char_p
    L.md_text_table.argtypes = [vp, C.c_int64, C.c_int, u64p, dp, C.c_char_p, C.c_int64, i64p]
    L.md_csv_write_table.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int64, C.c_int, u64p, dp]
    _lib = L
    return L
