"""ctypes binding of libfaco_hip.so, the C ABI in include/faco_engine.h.

The library is built in-tree (gplsaco_amd/libfaco_hip.so) by
gplsaco_amd.build.build(); importing this module never builds and never falls
back to Python: a missing library raises ImportError, so a product path can
not silently run on anything but the HIP engine.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfaco_hip.so"

FACO_OK, FACO_INVALID_ARGUMENT, FACO_RUNTIME_ERROR = 0, 1, 2

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
vp = ctypes.c_void_p


class Settings(ctypes.Structure):
    """faco_settings: the reference FixedFacoSettings layout (64 bytes)."""

    _fields_ = [
        ("ants", ctypes.c_uint32),
        ("primary_width", ctypes.c_uint32),
        ("backup_width", ctypes.c_uint32),
        ("ls_width", ctypes.c_uint32),
        ("beta", ctypes.c_double),
        ("retention", ctypes.c_double),
        ("p_best", ctypes.c_double),
        ("epoch_source_probability", ctypes.c_double),
        ("ls_evaluation_limit", ctypes.c_uint64),
        ("initial_ls_evaluation_limit", ctypes.c_uint64),
    ]


class AntInfo(ctypes.Structure):
    _fields_ = [
        ("construction_cost", ctypes.c_int64),
        ("final_cost", ctypes.c_int64),
        ("mne", ctypes.c_uint32),
        ("steps", ctypes.c_uint32),
        ("nonidentity_relocations", ctypes.c_uint32),
        ("legal_exhausted", ctypes.c_uint32),
        ("processed_nodes", ctypes.c_uint32),
        ("candidate_checks", ctypes.c_uint32),
        ("move_evaluations", ctypes.c_uint32),
        ("accepted_moves", ctypes.c_uint32),
        ("reactivations", ctypes.c_uint32),
        ("evaluation_limit_reached", ctypes.c_uint32),
        ("checklist_size", ctypes.c_uint32),
        ("start_node", ctypes.c_uint32),
        ("constraint_rejections", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


ANT_INFO_DTYPE = np.dtype([(name, np.int64 if t is ctypes.c_int64 else np.uint32) for name, t in AntInfo._fields_])
assert ANT_INFO_DTYPE.itemsize == ctypes.sizeof(AntInfo)

STAGE_COUNT = 12


class Result(ctypes.Structure):
    _fields_ = [
        ("evaluation_limit_per_colony", ctypes.c_uint64),
        ("completed_tour_evaluations_per_colony", ctypes.c_uint64),
        ("total_tour_evaluations", ctypes.c_uint64),
        ("launched_batches", ctypes.c_uint64),
        ("completed_batches", ctypes.c_uint64),
        ("completed_construction_steps", ctypes.c_uint64),
        ("completed_ls_evaluations", ctypes.c_uint64),
        ("allocated_device_bytes", ctypes.c_uint64),
        ("actual_seconds", ctypes.c_double),
        ("search_seconds", ctypes.c_double),
        ("search_gpu_milliseconds", ctypes.c_double),
        ("kernel_milliseconds", ctypes.c_double * STAGE_COUNT),
        ("completed_constraint_rejections", ctypes.c_uint64),
        ("persistent_launches", ctypes.c_uint64),
    ]


class Program(ctypes.Structure):
    _fields_ = [
        ("ir_version", ctypes.c_uint16),
        ("numeric_spec_id", ctypes.c_uint16),
        ("feature_spec_id", ctypes.c_uint16),
        ("length", ctypes.c_uint16),
        ("constants_count", ctypes.c_uint16),
        ("opcode", ctypes.c_uint8 * 63),
        ("operand", ctypes.c_uint8 * 63),
        ("constants", ctypes.c_float * 63),
    ]


class Controller(ctypes.Structure):
    """faco_controller: the reference ControllerProgram layout (1168 bytes)."""

    _fields_ = [("kind", ctypes.c_uint32), ("trees", Program * 3)]


class BaselinePolicy(ctypes.Structure):
    """faco_baseline_policy: the reference BaselinePolicy layout (48 bytes)."""

    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("mne_level", ctypes.c_uint32),
        ("max_mne_level", ctypes.c_uint32),
        ("region", ctypes.c_uint32),
        ("restart_mode", ctypes.c_uint32),
        ("restart_period", ctypes.c_uint32),
        ("restart_probability", ctypes.c_double),
        ("stagnation_step", ctypes.c_uint32),
        ("restart_stagnation", ctypes.c_uint32),
        ("restart_cooldown", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


class ControlState(ctypes.Structure):
    _fields_ = [
        ("archive_size", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("stagnant_batches", ctypes.c_uint64),
        ("epoch_batches", ctypes.c_uint64),
        ("restarts", ctypes.c_uint64),
        ("return_rate", ctypes.c_float),
        ("ls_work", ctypes.c_float),
    ]


class DeadlineSpec(ctypes.Structure):
    """faco_deadline_spec: a time-budget evaluation."""

    _fields_ = [
        ("budget_seconds", ctypes.c_double),
        ("mne_target", ctypes.c_uint32),
        ("preparation_mode", ctypes.c_uint32),
        ("program", ctypes.c_void_p),
        ("experiment_mask", ctypes.c_uint32),
        ("fixed_batches", ctypes.c_uint32),
        ("fixed_elapsed_ratio", ctypes.c_double),
        ("profile", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


class DeadlineResult(ctypes.Structure):
    _fields_ = [
        ("budget_seconds", ctypes.c_double),
        ("elapsed_seconds", ctypes.c_double),
        ("actual_seconds", ctypes.c_double),
        ("charged_seconds", ctypes.c_double),
        ("last_batch_completed_seconds", ctypes.c_double),
        ("overrun_seconds", ctypes.c_double),
        ("launched_batches", ctypes.c_uint64),
        ("completed_batches", ctypes.c_uint64),
        ("discarded_batches", ctypes.c_uint64),
        ("completed_construction_steps", ctypes.c_uint64),
        ("completed_ls_evaluations", ctypes.c_uint64),
        ("allocated_device_bytes", ctypes.c_uint64),
        ("preparation_completed", ctypes.c_uint32),
        ("profiled_batches", ctypes.c_uint32),
        ("setup_seconds", ctypes.c_double),
        ("upload_seconds", ctypes.c_double),
        ("initialization_gpu_milliseconds", ctypes.c_double),
    ]


class BatchProfile(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_uint32),
        ("committed", ctypes.c_uint32),
        ("wall_seconds", ctypes.c_double),
        ("download_seconds", ctypes.c_double),
        ("verification_seconds", ctypes.c_double),
        ("collection_seconds", ctypes.c_double),
        ("gpu_milliseconds", ctypes.c_double * 9),
    ]


CONTROL_PROGRAM, CONTROL_CONTROLLER, CONTROL_BASELINE, CONTROL_FACTORIAL = 1, 2, 3, 4


class ControlSpec(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_uint32),
        ("deciders", ctypes.c_uint32),
        ("programs", ctypes.c_void_p),
        ("controllers", ctypes.c_void_p),
        ("baseline", BaselinePolicy),
        ("factorial_variant", ctypes.c_uint32),
        ("experiment_mask", ctypes.c_uint32),
        ("force_fingerprint_collisions", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


class Checkpoints(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_void_p),
        ("count", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("tours", ctypes.c_void_p),
    ]


class ControlTrace(ctypes.Structure):
    _fields_ = [
        ("batches", ctypes.c_void_p),
        ("count", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("features", ctypes.c_void_p),
        ("scores", ctypes.c_void_p),
        ("masks", ctypes.c_void_p),
        ("actions", ctypes.c_void_p),
        ("alternatives", ctypes.c_void_p),
        ("regions", ctypes.c_void_p),
        ("costs", ctypes.c_void_p),
    ]


class Fork(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("enabled", ctypes.c_uint32),
        ("region", ctypes.c_uint32),
        ("mne", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


class TuneEntry(ctypes.Structure):
    _fields_ = [
        ("waves", ctypes.c_uint32),
        ("nbr_in_lds", ctypes.c_uint32),
        ("blocks_per_colony", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("ms_per_iteration", ctypes.c_double),
    ]


assert ctypes.sizeof(Settings) == 64
assert ctypes.sizeof(Program) == 388
assert ctypes.sizeof(Controller) == 1168
assert ctypes.sizeof(BaselinePolicy) == 48
assert ctypes.sizeof(ControlState) == 40
assert ctypes.sizeof(Fork) == 24

# (name, restype, argtypes) of every exported entry point of include/faco_engine.h
SIGNATURES = {
    "faco_last_error": (ctypes.c_char_p, []),
    "faco_device_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_size_t,
                                        ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "faco_prepare_common_initial": (ctypes.c_int, [f64p, ctypes.c_uint32, ctypes.POINTER(Settings),
                                                   ctypes.POINTER(Settings), u32p, ctypes.POINTER(ctypes.c_int64),
                                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                   vp, vp, vp, vp]),
    "faco_engine_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Settings), ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.POINTER(vp)]),
    "faco_engine_create_mixed": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Settings),
                                                ctypes.c_uint32, ctypes.POINTER(vp)]),
    "faco_engine_autotune": (ctypes.c_int, [vp, u64p, u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(TuneEntry),
                                            ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    "faco_engine_autotune_control": (ctypes.c_int, [vp, u64p, u64p, ctypes.c_uint32, ctypes.POINTER(ControlSpec),
                                                    ctypes.c_uint32, ctypes.POINTER(TuneEntry), ctypes.c_uint32,
                                                    ctypes.POINTER(ctypes.c_uint32)]),
    "faco_engine_destroy": (None, [vp]),
    "faco_engine_settings": (ctypes.c_int, [vp, ctypes.POINTER(Settings)]),
    "faco_engine_launch_shape": (ctypes.c_int, [vp, ctypes.c_uint32] + [ctypes.POINTER(ctypes.c_uint32)] * 5),
    "faco_engine_set_waves": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32]),
    "faco_engine_set_local_search": (ctypes.c_int, [vp, ctypes.c_uint32]),
    "faco_engine_set_roulette_margin": (ctypes.c_int, [vp, ctypes.c_double]),
    "faco_fixed_set_local_search": (ctypes.c_int, [vp, ctypes.c_uint32]),
    "faco_engine_register_problem": (ctypes.c_int, [vp, ctypes.c_uint64, f64p, ctypes.c_uint32,
                                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "faco_engine_preparation_profile": (ctypes.c_int, [vp, ctypes.c_uint64, f64p]),
    "faco_engine_problem_data": (ctypes.c_int, [vp, ctypes.c_uint64, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_int64)]),
    "faco_engine_evaluate_faco": (ctypes.c_int, [vp, u64p, u64p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                                 ctypes.c_uint32, ctypes.POINTER(Result), u32p, i64p, u64p,
                                                 ctypes.POINTER(Checkpoints)]),
    "faco_fixed_create": (ctypes.c_int, [f64p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(Settings),
                                         ctypes.POINTER(vp)]),
    "faco_fixed_destroy": (None, [vp]),
    "faco_fixed_settings": (ctypes.c_int, [vp, ctypes.POINTER(Settings)]),
    "faco_fixed_run": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_double), vp, vp, vp, vp, vp, vp, vp, vp]),
    "faco_diagnostic_construct_ls": (ctypes.c_int, [f64p, ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                                    u32p, u32p, u32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                                    u32p, u32p, u32p, u32p, vp, ctypes.c_uint32]),
    "faco_diagnostic_distances": (ctypes.c_int, [i32p, ctypes.c_uint32, i64p]),
    "faco_diagnostic_start_nodes": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                                   ctypes.c_uint32, ctypes.c_uint32, u32p]),
    "faco_diagnostic_copy": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "faco_diagnostic_escape_rows": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p,
                                                   ctypes.c_uint32, u32p, ctypes.c_uint64, ctypes.c_uint32, u32p,
                                                   ctypes.c_uint32, vp, u32p, u8p, u32p]),
    "faco_diagnostic_escape_transactions": (ctypes.c_int, [ctypes.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p, u64p,
                                                           u32p, u32p, ctypes.c_uint32, u32p, u32p, u32p, u32p]),
    "faco_diagnostic_escape_draws": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                    ctypes.c_uint32, ctypes.c_uint32, u32p]),
    "faco_diagnostic_philox": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_uint32, u32p]),
    "faco_score_cpu": (ctypes.c_int, [ctypes.POINTER(Program), f32p, u32p, ctypes.c_uint32, f32p, i32p]),
    "faco_score_gpu": (ctypes.c_int, [ctypes.POINTER(Program), f32p, u32p, ctypes.c_uint32, f32p, i32p]),
    "faco_score_controller_cpu": (ctypes.c_int, [ctypes.POINTER(Controller), f32p, u32p, ctypes.c_uint32, f32p,
                                                 i32p]),
    "faco_score_controller_gpu": (ctypes.c_int, [ctypes.POINTER(Controller), f32p, u32p, ctypes.c_uint32, f32p,
                                                 i32p]),
    "faco_prior_param_count": (ctypes.c_uint32, []),
    "faco_prior_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(vp)]),
    "faco_prior_destroy": (None, [vp]),
    "faco_prior_get_params": (ctypes.c_int, [vp, f32p]),
    "faco_prior_set_params": (ctypes.c_int, [vp, f32p]),
    "faco_prior_logits": (ctypes.c_int, [vp, f32p, ctypes.c_uint32, f32p]),
    "faco_prior_gradients": (ctypes.c_int, [vp, f32p, u8p, ctypes.c_uint32, ctypes.c_float, f32p,
                                            ctypes.POINTER(ctypes.c_float)]),
    "faco_prior_gradients_noalloc": (ctypes.c_int, [vp, f32p, u8p, ctypes.c_uint32, ctypes.c_float, f32p,
                                                    ctypes.POINTER(ctypes.c_float)]),
    "faco_prior_train": (ctypes.c_int, [vp, f32p, u8p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_float, ctypes.c_float, ctypes.c_uint64, f32p]),
    "faco_prior_mma_check": (ctypes.c_int, [f32p, f32p, f32p]),
    "faco_engine_edge_features": (ctypes.c_int, [vp, ctypes.c_uint64, f32p]),
    "faco_engine_set_edge_prior": (ctypes.c_int, [vp, ctypes.c_uint64, vp]),
    "faco_engine_apply_edge_prior": (ctypes.c_int, [vp, vp, ctypes.c_uint64, vp, ctypes.c_double]),
    "faco_engine_evaluate_deadline": (ctypes.c_int, [vp, u64p, u64p, ctypes.c_uint32, ctypes.POINTER(DeadlineSpec),
                                                     ctypes.POINTER(DeadlineResult), u32p, f64p, u8p, f64p, vp, vp,
                                                     ctypes.c_uint32, vp]),
    "faco_engine_set_preparation_charges": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]),
    "faco_engine_constraint_rejections": (ctypes.c_int, [vp, u64p, ctypes.c_uint32]),
    "faco_engine_escape_counters": (ctypes.c_int, [vp, u64p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    "faco_engine_evaluate_control": (ctypes.c_int, [vp, u64p, u64p, ctypes.c_uint32, ctypes.c_uint64,
                                                    ctypes.POINTER(ControlSpec), ctypes.c_uint32,
                                                    ctypes.POINTER(ControlTrace), ctypes.POINTER(Result), u32p, i64p,
                                                    u64p, vp, ctypes.POINTER(Checkpoints)]),
    "faco_engine_register_graph_problem": (ctypes.c_int, [vp, ctypes.c_uint64, f64p, ctypes.c_uint32, vp, ctypes.c_uint64,
                                                          u32p, u32p, vp, u32p, ctypes.POINTER(ctypes.c_double),
                                                          ctypes.POINTER(ctypes.c_double)]),
    "faco_engine_register_problems_device": (ctypes.c_int, [vp, u64p, f64p, vp, ctypes.c_uint32,
                                                            ctypes.POINTER(ctypes.c_double)]),
    "faco_state_from_bytes": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "faco_state_to_bytes": (ctypes.c_int, [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "faco_state_describe": (ctypes.c_int, [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "faco_state_select_colony": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]),
    "faco_state_destroy": (None, [vp]),
    "faco_engine_capture_control": (ctypes.c_int, [vp, u64p, u64p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                                   ctypes.POINTER(ControlSpec), ctypes.c_uint32,
                                                   ctypes.POINTER(ControlTrace), ctypes.POINTER(Result), u32p, i64p,
                                                   u64p, vp, ctypes.POINTER(ctypes.c_void_p)]),
    "faco_engine_continue_control": (ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.POINTER(ControlSpec),
                                                    ctypes.POINTER(Fork), ctypes.c_uint32, ctypes.POINTER(ControlTrace),
                                                    ctypes.POINTER(Result), u32p, i64p, u64p, vp]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libfaco_hip.so; raise ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run gplsaco_amd.build.build() (hipcc, gfx950) first")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = L
    return _lib


def check(status: int) -> None:
    """Map a C status to the reference's Python exceptions."""
    if status == FACO_OK:
        return
    message = lib().faco_last_error().decode(errors="replace")
    if status == FACO_INVALID_ARGUMENT:
        raise ValueError(message)
    raise RuntimeError(message)


def ptr(array) -> ctypes.c_void_p | None:
    return None if array is None else array.ctypes.data_as(ctypes.c_void_p)
