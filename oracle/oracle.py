"""ctypes loader for the CPU oracle (liboracle.so) and the reference checker.

TEST INFRASTRUCTURE ONLY. The oracle is a CPU restatement of the reference
FACO algorithm; it is the checker for the HIP engine and the CPU baseline of
bench.py. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this module. The product package (gplsaco_amd) never imports it.

Parity status of each oracle function and what pins it:

- euc2d: closed form (integer isqrt) checked against floor(sqrt(s)+0.5) in
  FP64 on exhaustive small ranges and random samples, and against hand values.
- philox4x32_10: Random123 known-answer vectors.
- uniform53 / start_node / colony_key: restated from the reference's domain
  split (cuda/faco_choices.cuh:32-84, cuda/batch_engine.cu:27-32,762); checked
  for the closed-form word layout, rejection threshold and domain separation.
  The raw cuRAND stream is not reproduced (parity unpinned at that boundary;
  the engine and the oracle share this Philox instead).
- prepare: k-NN rows by (distance, id) and NN tour checked by brute force;
  initial checklist 2-opt checked against the reference's checklist_two_opt
  (oracle/_ref).
- construct_ls: against the reference's FocusedConstruction +
  checklist_two_opt (oracle/_ref) on explicit visit orders, including the
  three-node identity-move case and the evaluation caps {0,1,7,127,inf}.
- select_next: against the reference's select_next (oracle/_ref), including
  the zero-weight, backup and global-fallback branches.
- trail limits / pheromone: against candidate_trail_limits and
  SparsePheromone (oracle/_ref).
- run_faco: composition of the above; the per-batch selection rule (argmin
  first index, strict improvement, epoch-source coin) restated from
  cuda/colony_state.cuh:148-208 and checked by replaying its trace through
  the parts above in tests/test_oracle.py.
- control (tour identity, canonical order, sampling, regions, archive update,
  action snapshot, feedback, baseline and factorial policies): against the
  reference's control_cpu.cpp / control_ops.hpp / baseline_policy.hpp /
  factorial_policy.hpp (oracle/_ref) on random states, and the hand-derived
  cases of tests/cpp/test_control_cpu.cpp and test_baseline_policy.cpp.
- controller scoring (conditional three-tree): restated from
  cuda/gp_score.cuh:37-62; pinned by closed forms (a kind-0 controller equals
  score_cpu; each stage is an argmax over the legal blocks) in
  tests/test_oracle_control.py. The reference has no CPU scorer for it.
- run_controlled: composition of the above in the per-batch order of
  cuda/batch_engine.cu:880-1060; its restart transaction
  (cuda/control_state.cuh:309-364) and archive/feedback update (:105-198)
  have no CPU counterpart in the reference and are pinned by the invariants
  in tests/test_oracle_control.py (archive = archive_update of the batch's
  pool, restarts reset the trails to the new upper bound, feedback = the
  reference's update_feedback).
- Hard mode (run_faco / run_controlled with graphs=): move predicates against
  complete before/after edge sets (tests/cpp/test_hard_edges.cpp's method),
  the Hard 2-opt and selection live against the reference core (oracle/_ref)
  in tests/test_oracle_hard.py.
- escape primitives (slot views, edge transactions, streams): the reference's
  scripted slot cases, exhaustive move checks and stream separation
  (tests/test_oracle_escape.py).
- Escape mode (escape_ant, run_faco / run_controlled with escape=): the
  reference's per-ant replay checks of tests/cpp/test_escape_engine.cpp
  (moves rebuild the tour, new edges only through permitted moves and within
  the cache, anchors = the five endpoints, LS views keep their valid counts,
  order and replacement bound, footprints mark replaced rows) and
  disabled-escape runs equal to Hard runs (tests/test_oracle_escape_mode.py).
  The escape stream is this Philox's, not cuRAND's (unpinned at that
  boundary, as above).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libgpfaco_ref.so"
REFERENCE_ROOT = Path("/root/reference")

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Settings(ctypes.Structure):
    """Mirror of the reference FixedFacoSettings (64 bytes)."""

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


class Feedback(ctypes.Structure):
    """FeedbackState (reference cpp/include/gp_faco/control_ops.hpp:160-163)."""

    _fields_ = [("return_rate", ctypes.c_float), ("ls_work", ctypes.c_float), ("stagnant_batches", ctypes.c_uint64),
                ("epoch_batches", ctypes.c_uint64), ("restarts", ctypes.c_uint64)]


class Baseline(ctypes.Structure):
    """BaselinePolicy (reference cpp/include/gp_faco/baseline_policy.hpp:20-27), 48 bytes.

    kind 0 static / 1 rule; restart_mode 0 none / 1 periodic / 2 bernoulli."""

    _fields_ = [("kind", ctypes.c_uint32), ("mne_level", ctypes.c_uint32), ("max_mne_level", ctypes.c_uint32),
                ("region", ctypes.c_uint32), ("restart_mode", ctypes.c_uint32), ("restart_period", ctypes.c_uint32),
                ("restart_probability", ctypes.c_double), ("stagnation_step", ctypes.c_uint32),
                ("restart_stagnation", ctypes.c_uint32), ("restart_cooldown", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class ControlOut(ctypes.Structure):
    _fields_ = [("archive_size", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("stagnant_batches", ctypes.c_uint64),
                ("epoch_batches", ctypes.c_uint64), ("restarts", ctypes.c_uint64), ("return_rate", ctypes.c_float),
                ("ls_work", ctypes.c_float)]


class Program(ctypes.Structure):
    """Program IR v1 record (reference cpp/include/gp_faco/program.hpp:10-19), 388 bytes."""

    _fields_ = [("ir_version", ctypes.c_uint16), ("numeric_spec_id", ctypes.c_uint16),
                ("feature_spec_id", ctypes.c_uint16), ("length", ctypes.c_uint16),
                ("constants_count", ctypes.c_uint16), ("opcode", ctypes.c_uint8 * 63),
                ("operand", ctypes.c_uint8 * 63), ("constants", ctypes.c_float * 63)]


class Controller(ctypes.Structure):
    """ControllerProgram (reference cpp/include/gp_faco/program.hpp:27-30), 1168 bytes."""

    _fields_ = [("kind", ctypes.c_uint32), ("trees", Program * 3)]


def make_program(opcode, operand, constants=(), feature_spec_id=2) -> Program:
    p = Program()
    p.ir_version, p.numeric_spec_id, p.feature_spec_id = 1, 1, feature_spec_id
    p.length, p.constants_count = len(opcode), len(constants)
    for i, (o, a) in enumerate(zip(opcode, operand)):
        p.opcode[i], p.operand[i] = o, a
    for i, c in enumerate(constants):
        p.constants[i] = c
    return p


def default_settings(**overrides) -> Settings:
    s = Settings(0, 16, 64, 20, 1.0, 0.5, 0.1, 0.01, 100000, 100000)
    for key, value in overrides.items():
        setattr(s, key, value)
    return s


def build(reference: bool | None = None) -> None:
    """Compile liboracle.so (and the reference checker when the checkout exists)."""
    targets = ["all"]
    if reference or (reference is None and REFERENCE_ROOT.exists()):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", str(HERE), *targets], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build(reference=False)
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_euc2d.restype = ctypes.c_int64
        L.oracle_euc2d.argtypes = [ctypes.c_int32] * 4
        L.oracle_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.oracle_set_alternate_stream.argtypes = [ctypes.c_int]
        L.oracle_set_alternate_stream.restype = None
        L.oracle_uniform53.restype = ctypes.c_double
        L.oracle_uniform53.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_start_node.restype = ctypes.c_uint32
        L.oracle_start_node.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_colony_key.restype = ctypes.c_uint64
        L.oracle_colony_key.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_mixed.restype = ctypes.c_uint64
        L.oracle_mixed.argtypes = [ctypes.c_uint64]
        L.oracle_trail_limits.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_prepare.argtypes = [ctypes.c_uint32, i32p, ctypes.POINTER(Settings), ctypes.POINTER(Settings),
                                     u32p, u32p, u32p, f64p, u32p, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64)]
        L.oracle_construct_ls.argtypes = [ctypes.c_uint32, i32p, u32p, ctypes.c_uint32, u32p, u32p,
                                          ctypes.c_uint32, ctypes.c_uint64, u32p, u32p, u32p, u32p, i64p, u64p,
                                          ctypes.c_uint32]
        L.oracle_select_next.argtypes = [ctypes.c_uint32, i32p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p,
                                         f64p, u8p, ctypes.c_uint32, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int)]
        L.oracle_run_faco.argtypes = [ctypes.c_uint32, ctypes.c_uint32, i32p, u64p, ctypes.c_uint32, u64p,
                                      ctypes.POINTER(Settings), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                      u32p, i64p, u64p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.oracle_escape_ant.argtypes = [ctypes.c_uint32, i32p, ctypes.POINTER(Settings), ctypes.c_void_p, u32p, u32p,
                                        u32p, f64p, ctypes.c_double, u8p, ctypes.c_uint64, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p,
                                        u8p, u32p, u8p, u8p, u32p, ctypes.POINTER(ctypes.c_uint32), u32p,
                                        ctypes.POINTER(ctypes.c_uint32), u64p, u64p, i64p]
        L.oracle_prior_features.argtypes = [ctypes.c_uint32, i32p, ctypes.c_uint32, u32p, f64p, ctypes.c_double, f32p]
        L.oracle_escape_slot_view.argtypes = [u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, u32p,
                                              u8p, u32p]
        L.oracle_escape_draws.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32, u32p]
        L.oracle_escape_transaction.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, u32p, ctypes.c_uint32, u32p,
                                                ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, u32p, u32p, u32p, u32p]
        L.oracle_hard_relocation_allowed.argtypes = [u32p, u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                     u32p, u32p]
        L.oracle_select_next_hard.argtypes = [ctypes.c_uint32, i32p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, f64p,
                                              u8p, ctypes.c_uint32, ctypes.c_double, u32p, u32p, u32p,
                                              ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int)]
        L.oracle_hard_two_opt_allowed.argtypes = [ctypes.c_uint32] * 4 + [u32p, u32p, ctypes.c_uint32]
        L.oracle_escape_relocation.argtypes = [u32p, u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                               u32p, u32p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, u32p]
        L.oracle_escape_commit.argtypes = [u32p, u32p, ctypes.c_uint32, ctypes.c_uint64, u32p, u32p, u32p]
        L.oracle_max_threads.restype = ctypes.c_int
        L.oracle_tour_ident.argtypes = [ctypes.c_uint32, u32p, u64p]
        L.oracle_canonical_compare.restype = ctypes.c_int
        L.oracle_canonical_compare.argtypes = [ctypes.c_uint32, u32p, u32p]
        L.oracle_sample_nodes.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p]
        L.oracle_remaining_node.restype = ctypes.c_uint32
        L.oracle_remaining_node.argtypes = [ctypes.c_uint32, u32p, ctypes.c_uint32]
        L.oracle_build_region.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_uint64,
                                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p]
        L.oracle_baseline_action.restype = ctypes.c_uint32
        L.oracle_baseline_action.argtypes = [ctypes.POINTER(Baseline), ctypes.POINTER(Feedback), ctypes.c_uint32,
                                             ctypes.c_double, ctypes.c_uint32]
        L.oracle_baseline_uniform.restype = ctypes.c_double
        L.oracle_baseline_uniform.argtypes = [ctypes.POINTER(Baseline), ctypes.c_uint64, ctypes.c_uint32]
        L.oracle_factorial_mask.restype = ctypes.c_uint32
        L.oracle_factorial_mask.argtypes = [ctypes.c_uint32, ctypes.POINTER(Baseline), ctypes.POINTER(Feedback),
                                            ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32]
        L.oracle_update_feedback.argtypes = [ctypes.POINTER(Feedback), ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_uint32]
        L.oracle_restart_feedback.argtypes = [ctypes.POINTER(Feedback)]
        L.oracle_score_controller.argtypes = [ctypes.c_void_p, f32p, ctypes.c_uint32, f32p,
                                              ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_archive_update.argtypes = [ctypes.c_uint32, i32p, u32p, ctypes.c_uint32, u32p, ctypes.c_uint32,
                                            u32p, u32p, ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_action_snapshot.argtypes = [ctypes.c_uint32, i32p, ctypes.c_uint32, u32p, f64p, ctypes.c_double,
                                             u32p, u32p, ctypes.c_uint32, u32p, f64p, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.POINTER(Feedback),
                                             ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                             f32p, u32p, ctypes.POINTER(ctypes.c_uint32),
                                             ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_run_controlled.argtypes = [ctypes.c_uint32, ctypes.c_uint32, i32p, u64p, ctypes.c_uint32, u64p,
                                            ctypes.POINTER(Settings), ctypes.c_uint64, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.POINTER(Baseline), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                            ctypes.c_int,
                                            u32p, i64p, u64p, ctypes.c_void_p] + [ctypes.c_void_p] * 7 + [
                                            ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.oracle_colony_select.argtypes = [ctypes.c_uint32, i32p, ctypes.POINTER(Settings), ctypes.c_uint64,
                                           ctypes.c_uint32, ctypes.c_uint32, u32p, i64p, i64p, u32p, f64p, f64p,
                                           f64p, f64p, ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)]
        L.oracle_control_restart.argtypes = [ctypes.c_uint32, i32p, ctypes.POINTER(Settings), ctypes.c_uint32, u32p,
                                             i64p, ctypes.c_uint32, ctypes.c_int, i64p, u32p, f64p, f64p, f64p, f64p,
                                             ctypes.POINTER(Feedback), u64p]
        L.oracle_two_opt.argtypes = [ctypes.c_uint32, i32p, u32p, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32,
                                     ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.c_uint32]
        L.oracle_two_opt_hard.argtypes = [ctypes.c_uint32, i32p, u32p, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32,
                                          ctypes.c_uint64, u32p, u32p, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
        _lib = L
    return _lib


def _check(status, library):
    if status == 1:
        raise ValueError(library.oracle_last_error().decode() if hasattr(library, "oracle_last_error")
                         else library.ref_last_error().decode())
    if status:
        raise RuntimeError(library.oracle_last_error().decode() if hasattr(library, "oracle_last_error")
                           else library.ref_last_error().decode())


def euc2d(x1, y1, x2, y2) -> int:
    return int(lib().oracle_euc2d(x1, y1, x2, y2))


def distance_matrix(xy: np.ndarray) -> np.ndarray:
    """EUC_2D int64 matrix from integer coordinates, via the oracle's rule."""
    xy = np.ascontiguousarray(xy, dtype=np.int64)
    diff = xy[:, None, :] - xy[None, :, :]
    s = (diff * diff).sum(-1).astype(np.uint64)
    r = np.floor(np.sqrt(s.astype(np.float64))).astype(np.uint64)
    r = np.where(r * r > s, r - 1, r)
    r = np.where((r + 1) * (r + 1) <= s, r + 1, r)
    return (r + (s > r * r + r)).astype(np.int64)


def philox(counter, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(np.asarray(counter, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def uniform53(key, batch, ant, step) -> float:
    return float(lib().oracle_uniform53(key, batch, ant, step))


def start_node(key, batch, ant, count) -> int:
    return int(lib().oracle_start_node(key, batch, ant, count))


def colony_key(instance_key, seed) -> int:
    return int(lib().oracle_colony_key(instance_key, seed))


def mixed(value) -> int:
    return int(lib().oracle_mixed(value))


def trail_limits(width, p_best, retention, cost):
    lo, hi = ctypes.c_double(), ctypes.c_double()
    lib().oracle_trail_limits(width, p_best, retention, cost, ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def prepare(xy: np.ndarray, settings: Settings | None = None) -> dict:
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    s = settings or default_settings()
    norm = Settings()
    # widths are clamped by the oracle; allocate for the requested maxima
    pw, bw, lw = s.primary_width, s.backup_width, s.ls_width
    primary = np.zeros(n * max(pw, 1), dtype=np.uint32)
    backup = np.zeros(n * max(bw, 1), dtype=np.uint32)
    ls = np.zeros(n * max(lw, 1), dtype=np.uint32)
    scale = np.zeros(n, dtype=np.float64)
    initial = np.zeros(n, dtype=np.uint32)
    cost, cheap = ctypes.c_int64(), ctypes.c_int64()
    L = lib()
    _check(L.oracle_prepare(n, xy, ctypes.byref(s), ctypes.byref(norm), primary, backup, ls, scale,
                            initial, ctypes.byref(cost), ctypes.byref(cheap)), L)
    return {
        "settings": norm,
        "primary": primary[: n * norm.primary_width].reshape(n, norm.primary_width),
        "backup": backup[: n * norm.backup_width].reshape(n, norm.backup_width),
        "ls": ls[: n * norm.ls_width].reshape(n, norm.ls_width),
        "local_scale": scale,
        "initial_tour": initial,
        "initial_cost": cost.value,
        "cheap_cost": cheap.value,
    }


COUNTER_NAMES = ("mne", "steps", "nonidentity_relocations", "legal_exhausted", "processed_nodes",
                 "candidate_checks", "move_evaluations", "accepted_moves", "reactivations",
                 "evaluation_limit_reached", "checklist_size")


def construct_ls(xy, ls, parent, visit, mne_target, limit, or_opt=False) -> dict:
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    ls = np.ascontiguousarray(ls, dtype=np.uint32)
    lw = ls.shape[1]
    out = {k: np.zeros(n, dtype=np.uint32) for k in ("construction_tour", "tour", "positions")}
    checklist = np.zeros(5 * n + 8, dtype=np.uint32)
    costs = np.zeros(2, dtype=np.int64)
    counters = np.zeros(11, dtype=np.uint64)
    L = lib()
    _check(L.oracle_construct_ls(n, xy, ls.reshape(-1), lw, np.ascontiguousarray(parent, dtype=np.uint32),
                                 np.ascontiguousarray(visit, dtype=np.uint32), mne_target, limit,
                                 out["construction_tour"], out["tour"], out["positions"], checklist, costs,
                                 counters, int(bool(or_opt))), L)
    out["counters"] = dict(zip(COUNTER_NAMES, (int(c) for c in counters)))
    out["checklist"] = checklist[: out["counters"]["checklist_size"]].copy()
    out["construction_cost"], out["final_cost"] = int(costs[0]), int(costs[1])
    return out


def two_opt(xy, ls, tour, checklist, limit, or_opt=False):
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    ls = np.ascontiguousarray(ls, dtype=np.uint32)
    tour = np.ascontiguousarray(tour, dtype=np.uint32).copy()
    checklist = np.ascontiguousarray(checklist, dtype=np.uint32)
    cost, evals = ctypes.c_int64(), ctypes.c_uint64()
    L = lib()
    _check(L.oracle_two_opt(n, xy, ls.reshape(-1), ls.shape[1], tour, checklist, checklist.size, limit,
                            ctypes.byref(cost), ctypes.byref(evals), int(bool(or_opt))), L)
    return tour, cost.value, evals.value


def two_opt_hard(xy, ls, tour, checklist, limit, graph):
    """Hard-mode checklist 2-opt inside `graph` (csr()): (tour, cost,
    evaluations, rejected moves)."""
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    ls = np.ascontiguousarray(ls, dtype=np.uint32)
    tour = np.ascontiguousarray(tour, dtype=np.uint32).copy()
    checklist = np.ascontiguousarray(checklist, dtype=np.uint32)
    offsets = np.ascontiguousarray(graph[0], dtype=np.uint32)
    neighbors = np.ascontiguousarray(graph[1], dtype=np.uint32)
    if not neighbors.size:
        neighbors = np.zeros(1, dtype=np.uint32)
    cost, evals, rejected = ctypes.c_int64(), ctypes.c_uint64(), ctypes.c_uint64()
    L = lib()
    _check(L.oracle_two_opt_hard(n, xy, ls.reshape(-1), ls.shape[1], tour, checklist, checklist.size, limit,
                                 offsets, neighbors, ctypes.byref(cost), ctypes.byref(evals),
                                 ctypes.byref(rejected)), L)
    return tour, cost.value, evals.value, rejected.value


def select_next(xy, primary_row, backup_row, products, visited, current, u):
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    chosen, stage = ctypes.c_uint32(), ctypes.c_int()
    primary_row = np.ascontiguousarray(primary_row, dtype=np.uint32)
    backup_row = np.ascontiguousarray(backup_row, dtype=np.uint32)
    L = lib()
    _check(L.oracle_select_next(n, xy, primary_row.size, backup_row.size, primary_row,
                                backup_row if backup_row.size else np.zeros(1, np.uint32),
                                np.ascontiguousarray(products, dtype=np.float64),
                                np.ascontiguousarray(visited, dtype=np.uint8), current, u,
                                ctypes.byref(chosen), ctypes.byref(stage)), L)
    return chosen.value, stage.value


def _graph_arrays(graphs, instances):
    """(offsets [instances][n+1], neighbours [instances][stride], stride) of
    one CSR graph per instance, or (None, None, 0)."""
    if graphs is None:
        return None, None, 0
    if len(graphs) != instances:
        raise ValueError("one graph per instance")
    stride = max(len(nb) for _, nb in graphs)
    offsets = np.ascontiguousarray(np.stack([np.asarray(o, dtype=np.uint32) for o, _ in graphs]))
    neighbors = np.zeros((instances, stride), dtype=np.uint32)
    for g, (_, nb) in enumerate(graphs):
        neighbors[g, :len(nb)] = nb
    return offsets, neighbors, stride


class alternate_stream:
    """Context manager: oracle runs inside draw every uniform from an
    unrelated generator (SplitMix64) instead of Philox, for statistical
    comparisons of the algorithm under independent random streams."""

    def __enter__(self):
        lib().oracle_set_alternate_stream(1)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_alternate_stream(0)
        return False


def _spec_rows(rows, instances):
    """One (primary, backup, ls) triple per instance ([n][width] each, right
    padded with n) as the oracle's contiguous block, or None."""
    if rows is None:
        return None
    if len(rows) != instances:
        raise ValueError("one set of rows per instance")
    return np.ascontiguousarray(np.stack([np.concatenate([np.asarray(r, dtype=np.uint32).reshape(-1) for r in triple])
                                          for triple in rows]))


ESCAPE_COUNTERS = ("construction_opportunities", "construction_gates", "construction_replaced_slots",
                   "escape_relocations", "ls_anchor_nodes", "ls_replaced_slots", "old_view_reactivations",
                   "anchor_reactivations", "new_edges", "construction_capacity_rejections", "ls_capacity_rejections")
ESCAPE_MODES = {None: 0, "on": 1, "disabled": 2}


def run_faco(xy_instances, instance_keys, seeds, settings, batches, mne_target=8, threads=0, trace=False,
             or_opt=False, multipliers=None, graphs=None, timed_from=0, rows=None, escape=None):
    """Fixed-iteration FACO over colonies (colony i uses instance i % instances).
    search_seconds times batches timed_from..batches-1 only (a window of the run).
    `multipliers` [instances, n, pw] (float32): learned heuristic multipliers,
    eta' = eta * multiplier in FP64. `graphs`: one undirected CSR
    (offsets, neighbors) per instance (csr()) for the Hard mode, in which
    every move must keep the tour inside its graph; `rows`: one (primary,
    backup, ls) triple per instance replacing the prepared rows, as a graph
    spec's own rows do. `escape` ("on" or "disabled", with `graphs`): the
    Escape mode on those graphs; "escape_counters" [colonies][11] in
    ESCAPE_COUNTERS order."""
    xy_instances = np.ascontiguousarray(xy_instances, dtype=np.int32)
    instances, n = xy_instances.shape[0], xy_instances.shape[1]
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    colonies = seeds.size
    keys = np.ascontiguousarray(instance_keys, dtype=np.uint64)
    tours = np.zeros((colonies, n), dtype=np.uint32)
    costs = np.zeros(colonies, dtype=np.int64)
    work = np.zeros((colonies, 2), dtype=np.uint64)
    rejections = np.zeros(colonies, dtype=np.uint64)
    escape_counters = np.zeros((colonies, len(ESCAPE_COUNTERS)), dtype=np.uint64)
    elapsed = ctypes.c_double()
    trace_parent = trace_trails = trace_costs = None
    if trace:
        norm = prepare(xy_instances[0], settings)["settings"]
        trace_parent = np.zeros((batches, n), dtype=np.uint32)
        trace_trails = np.zeros((batches, n * norm.primary_width), dtype=np.float64)
        trace_costs = np.zeros((batches, 3), dtype=np.int64)
    if multipliers is not None:
        mult = np.ascontiguousarray(multipliers, dtype=np.float32)
        if mult.shape[0] != instances:
            raise ValueError("one multiplier table per instance")
    g_offsets = g_neighbors = None
    stride = 0
    if graphs is not None:
        if len(graphs) != instances:
            raise ValueError("one graph per instance")
        stride = max(len(nb) for _, nb in graphs)
        g_offsets = np.ascontiguousarray(np.stack([np.asarray(o, dtype=np.uint32) for o, _ in graphs]))
        g_neighbors = np.zeros((instances, stride), dtype=np.uint32)
        for g, (_, nb) in enumerate(graphs):
            g_neighbors[g, :len(nb)] = nb
    spec_rows = _spec_rows(rows, instances)   # kept alive through the call
    L = lib()
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None  # noqa: E731
    _check(L.oracle_run_faco(n, instances, xy_instances.reshape(-1), keys, colonies, seeds,
                             ctypes.byref(settings), batches, mne_target, threads, tours.reshape(-1), costs,
                             work.reshape(-1), ptr(trace_parent), ptr(trace_trails), ptr(trace_costs),
                             ctypes.byref(elapsed), int(bool(or_opt)),
                             ptr(None if multipliers is None else mult), ptr(g_offsets), ptr(g_neighbors), stride,
                             int(timed_from), ptr(rejections), ptr(spec_rows), ESCAPE_MODES[escape],
                             ptr(escape_counters)), L)
    out = {"tours": tours, "costs": costs, "work": work, "search_seconds": elapsed.value,
           "constraint_rejections": rejections, "escape_counters": escape_counters}
    if trace:
        out.update(trace_parent=trace_parent, trace_trails=trace_trails, trace_costs=trace_costs)
    return out


def escape_ant(xy, settings, graph, parent, products, default_trail, parent_footprint, key, batch, ant, start,
               target, rows=None, disabled=False):
    """One Escape-mode ant on explicit colony state (parent tour, products
    [n][pw], default trail, parent footprint [n]) over the CSR `graph`,
    with its trace: "events" [moves][6] = (kind 0 relocation / 1 successor
    2-opt / 2 predecessor 2-opt, a, b, slot, cache size after, permit),
    "anchors", "ls_rows" [n][lw], "replaced", "footprint", "cache" [size][2],
    "counters" (ESCAPE_COUNTERS order), "ls" = evaluations, accepted,
    rejections, checklist size, "costs" = construction, final."""
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    norm = prepare(xy, settings)["settings"]
    spec_rows = _spec_rows(None if rows is None else [rows], 1)
    offsets, neighbors = (np.ascontiguousarray(a, dtype=np.uint32) for a in graph)
    out = {"construction_tour": np.zeros(n, np.uint32), "tour": np.zeros(n, np.uint32),
           "anchors": np.zeros(n, np.uint8), "ls_rows": np.zeros((n, norm.ls_width), np.uint32),
           "replaced": np.zeros((n, norm.ls_width), np.uint8), "footprint": np.zeros(n, np.uint8)}
    cache = np.zeros((64, 2), np.uint32)
    events = np.zeros((2 * n, 6), np.uint32)
    counters = np.zeros(len(ESCAPE_COUNTERS), np.uint64)
    ls = np.zeros(4, np.uint64)
    costs = np.zeros(2, np.int64)
    size, count = ctypes.c_uint32(), ctypes.c_uint32()
    L = lib()
    _check(L.oracle_escape_ant(
        n, xy, ctypes.byref(settings), spec_rows.ctypes.data_as(ctypes.c_void_p) if spec_rows is not None else None,
        offsets, neighbors, np.ascontiguousarray(parent, dtype=np.uint32),
        np.ascontiguousarray(products, dtype=np.float64).reshape(-1), float(default_trail),
        np.ascontiguousarray(parent_footprint, dtype=np.uint8), key, batch, ant, start, target, int(bool(disabled)),
        out["construction_tour"], out["tour"], out["anchors"], out["ls_rows"].reshape(-1), out["replaced"].reshape(-1),
        out["footprint"], cache.reshape(-1), ctypes.byref(size), events.reshape(-1), ctypes.byref(count), counters,
        ls, costs), L)
    out.update(cache=cache[:size.value], events=events[:count.value], counters=counters, ls=ls, costs=costs)
    return out


def prior_features(xy, primary, local_scale, epsilon):
    """Learned-prior edge features [n * pw, 8] (float32) of the candidate edges."""
    xy = np.ascontiguousarray(xy, dtype=np.int32)
    primary = np.ascontiguousarray(primary, dtype=np.uint32)
    n, pw = primary.shape
    scale = np.ascontiguousarray(local_scale, dtype=np.float64)
    out = np.zeros((n * pw, 8), dtype=np.float32)
    L = lib()
    _check(L.oracle_prior_features(n, xy.reshape(-1), pw, primary.reshape(-1), scale, float(epsilon),
                                   out.reshape(-1)), L)
    return out


# ----------------------------------------------------------------------------
# Escape primitives (cpp/include/gp_faco/escape_slots.hpp, escape_edges.hpp).

STATUS = {0: "allowed", 1: "invalid_edge", 2: "outside_graph", 3: "capacity"}


def escape_slot_view(original, backup, n, from_city, *, script=None, key=0, batch=0, ant=0, event=0, ls=False,
                     distance_of=None):
    """One slot view. script=r: the reference test's scripted draws (draw() = r,
    below(c) = c - 1); else the escape stream of (key, batch, ant, event, ls).
    distance_of [n] (float64): sort as an LS row by (distance_of[node], node)."""
    o = np.ascontiguousarray(original, dtype=np.uint32)
    b = np.ascontiguousarray(backup, dtype=np.uint32)
    members = np.zeros(o.size, dtype=np.uint32)
    replaced = np.zeros(o.size, dtype=np.uint8)
    counts = np.zeros(2, dtype=np.uint32)
    dist = None if distance_of is None else np.ascontiguousarray(distance_of, dtype=np.float64)
    L = lib()
    _check(L.oracle_escape_slot_view(o, o.size, b if b.size else np.zeros(1, np.uint32), b.size, n, from_city,
                                     0 if script is None else script + 1, key, batch, ant, event, 1 if ls else 0,
                                     None if dist is None else dist.ctypes.data_as(ctypes.c_void_p), members, replaced,
                                     counts), L)
    return {"members": members, "replaced": replaced, "valid": int(counts[0]), "replacements": int(counts[1])}


def escape_draws(key, batch, ant, event, ls, count):
    out = np.zeros(count, dtype=np.uint32)
    lib().oracle_escape_draws(key, batch, ant, event, 1 if ls else 0, count, out)
    return out


def _csr(graph_edges, n):
    rows = [[] for _ in range(n)]
    for a, b in graph_edges:
        rows[a].append(b)
        rows[b].append(a)
    offsets = np.zeros(n + 1, dtype=np.uint32)
    neighbors = []
    for a in range(n):
        r = sorted(set(rows[a]))
        neighbors += r
        offsets[a + 1] = offsets[a] + len(r)
    return offsets, np.array(neighbors or [0], dtype=np.uint32)


def csr(graph_edges, n):
    """Undirected CSR (offsets [n+1], sorted neighbours) of an edge list."""
    return _csr(graph_edges, n)


def hard_relocation_allowed(tour, target, node, csr_graph):
    """Hard-mode relocation predicate: `node` moved to follow `target`."""
    tour = np.ascontiguousarray(tour, dtype=np.uint32)
    n = tour.size
    positions = np.zeros(n, dtype=np.uint32)
    positions[tour] = np.arange(n, dtype=np.uint32)
    offsets, neighbors = csr_graph
    return bool(lib().oracle_hard_relocation_allowed(tour, positions, n, target, node, offsets, neighbors))


def select_next_hard(xy, primary_row, backup_row, products, visited, current, u, tour, csr_graph):
    """One Hard-mode roulette decision; returns (chosen, stage) with stage 0
    primary, 1 zero-weight, 2 backup, 3 global fallback, 4 exhausted."""
    xy = np.ascontiguousarray(xy, dtype=np.int32)
    n = xy.shape[0]
    chosen, stage = ctypes.c_uint32(), ctypes.c_int()
    offsets, neighbors = csr_graph
    _check(lib().oracle_select_next_hard(
        n, xy.reshape(-1), len(primary_row), len(backup_row), np.ascontiguousarray(primary_row, dtype=np.uint32),
        np.ascontiguousarray(backup_row, dtype=np.uint32), np.ascontiguousarray(products, dtype=np.float64),
        np.ascontiguousarray(visited, dtype=np.uint8), current, u, np.ascontiguousarray(tour, dtype=np.uint32),
        offsets, neighbors, ctypes.byref(chosen), ctypes.byref(stage)), lib())
    return chosen.value, stage.value


def hard_two_opt_allowed(a, a_neighbor, b, b_neighbor, csr_graph, n):
    """Hard-mode 2-opt predicate: (a, a_neighbor), (b, b_neighbor) -> (a, b), (a_neighbor, b_neighbor)."""
    offsets, neighbors = csr_graph
    return bool(lib().oracle_hard_two_opt_allowed(a, a_neighbor, b, b_neighbor, offsets, neighbors, n))


def _proposal(p):
    return {"status": STATUS[int(p[0])], "count": int(p[1]), "expected_size": int(p[2]),
            "expected_epoch": int(p[3]) | (int(p[4]) << 32),
            "edges": [(int(p[5 + 2 * i]), int(p[6 + 2 * i])) for i in range(int(p[1]))]}


def escape_transaction(removed, added, csr_graph, n, cache_edges, cache_epoch, permit):
    """prepare_escape_edges + commit_escape_edges on a copy of the cache."""
    r = np.ascontiguousarray(removed, dtype=np.uint32).reshape(-1, 2)
    a = np.ascontiguousarray(added, dtype=np.uint32).reshape(-1, 2)
    cache = np.zeros((64, 2), dtype=np.uint32)
    if len(cache_edges):
        cache[: len(cache_edges)] = np.asarray(cache_edges, dtype=np.uint32)
    offsets, neighbors = csr_graph
    prop = np.zeros(11, dtype=np.uint32)
    committed = np.zeros(1, dtype=np.uint32)
    after = np.zeros((64, 2), dtype=np.uint32)
    size_after = np.zeros(1, dtype=np.uint32)
    L = lib()
    _check(L.oracle_escape_transaction(r.shape[0], r.reshape(-1), a.reshape(-1), offsets, neighbors, n,
                                       cache.reshape(-1), len(cache_edges), cache_epoch, 1 if permit else 0, prop,
                                       committed, after.reshape(-1), size_after), L)
    out = _proposal(prop)
    out.update(committed=bool(committed[0]), cache_after=[tuple(e) for e in after[: int(size_after[0])].tolist()])
    return out


def escape_commit(proposal, cache_edges, cache_epoch):
    """commit_escape_edges of a proposal dict (from escape_transaction) into a
    copy of another cache state; returns (committed, cache edges after)."""
    raw = np.zeros(11, dtype=np.uint32)
    raw[0] = {v: k for k, v in STATUS.items()}[proposal["status"]]
    raw[1], raw[2] = proposal["count"], proposal["expected_size"]
    raw[3], raw[4] = proposal["expected_epoch"] & 0xFFFFFFFF, proposal["expected_epoch"] >> 32
    for i, (a, b) in enumerate(proposal["edges"]):
        raw[5 + 2 * i], raw[6 + 2 * i] = a, b
    cache = np.zeros((64, 2), dtype=np.uint32)
    if len(cache_edges):
        cache[: len(cache_edges)] = np.asarray(cache_edges, dtype=np.uint32)
    committed = np.zeros(1, dtype=np.uint32)
    after = np.zeros((64, 2), dtype=np.uint32)
    size_after = np.zeros(1, dtype=np.uint32)
    L = lib()
    _check(L.oracle_escape_commit(raw, cache.reshape(-1), len(cache_edges), cache_epoch, committed, after.reshape(-1),
                                  size_after), L)
    return bool(committed[0]), [tuple(e) for e in after[: int(size_after[0])].tolist()]


def escape_relocation(tour, target, node, csr_graph, cache_edges, cache_epoch, permit):
    """prepare_escape_relocation of moving `node` after `target` in `tour`."""
    t = np.ascontiguousarray(tour, dtype=np.uint32)
    n = t.size
    pos = np.zeros(n, dtype=np.uint32)
    pos[t] = np.arange(n, dtype=np.uint32)
    cache = np.zeros((64, 2), dtype=np.uint32)
    if len(cache_edges):
        cache[: len(cache_edges)] = np.asarray(cache_edges, dtype=np.uint32)
    offsets, neighbors = csr_graph
    prop = np.zeros(11, dtype=np.uint32)
    L = lib()
    _check(L.oracle_escape_relocation(t, pos, n, target, node, offsets, neighbors, cache.reshape(-1), len(cache_edges),
                                      cache_epoch, 1 if permit else 0, prop), L)
    return _proposal(prop)


def max_threads() -> int:
    return int(lib().oracle_max_threads())


# ----------------------------------------------------------------------------
# Reference checker (built from the reference's own sources, oracle/_ref).

def ref_available() -> bool:
    return REF_PATH.exists()


def ref():
    global _ref
    if _ref is None:
        if not REF_PATH.exists():
            raise FileNotFoundError("reference checker not built (oracle/_ref)")
        R = ctypes.CDLL(str(REF_PATH))
        R.ref_last_error.restype = ctypes.c_char_p
        R.ref_construct_ls.argtypes = [ctypes.c_uint32, f64p, u32p, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32,
                                       ctypes.c_uint64, u32p, u32p, u32p, u32p, f64p, u64p]
        R.ref_two_opt.argtypes = [ctypes.c_uint32, f64p, u32p, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32,
                                  ctypes.c_uint64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]
        R.ref_select_next.argtypes = [ctypes.c_uint32, f64p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, f64p,
                                      u8p, ctypes.c_uint32, ctypes.c_double, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_int)]
        R.ref_two_opt_hard.argtypes = [ctypes.c_uint32, f64p, u32p, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32,
                                       ctypes.c_uint64, u32p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double), u64p]
        R.ref_select_next_hard.argtypes = [ctypes.c_uint32, f64p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, f64p,
                                           u8p, ctypes.c_uint32, ctypes.c_double, u32p, u32p, ctypes.c_uint64,
                                           ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int)]
        R.ref_trail_limits.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        R.ref_pheromone_run.argtypes = [ctypes.c_uint32, f64p, u32p, ctypes.c_uint32, ctypes.c_double,
                                        ctypes.c_uint32, u32p, ctypes.c_double, f64p, f64p, f64p, ctypes.c_double,
                                        f64p, f64p, f64p]
        R.ref_score_cpu.argtypes = [u8p, u8p, ctypes.c_uint32, f32p, ctypes.c_uint32, ctypes.c_uint16, f32p,
                                    u32p, ctypes.c_uint32, f32p, ctypes.POINTER(ctypes.c_int32)]
        _ref = R
    return _ref


def ref_construct_ls(matrix, ls, parent, visit, mne_target, limit) -> dict:
    matrix = np.ascontiguousarray(matrix, dtype=np.float64)
    n = matrix.shape[0]
    ls = np.ascontiguousarray(ls, dtype=np.uint32)
    out = {k: np.zeros(n, dtype=np.uint32) for k in ("construction_tour", "tour", "positions")}
    checklist = np.zeros(5 * n + 8, dtype=np.uint32)
    costs = np.zeros(2, dtype=np.float64)
    counters = np.zeros(11, dtype=np.uint64)
    R = ref()
    _check(R.ref_construct_ls(n, matrix.reshape(-1), ls.reshape(-1), ls.shape[1],
                              np.ascontiguousarray(parent, dtype=np.uint32),
                              np.ascontiguousarray(visit, dtype=np.uint32), mne_target, limit,
                              out["construction_tour"], out["tour"], out["positions"], checklist, costs,
                              counters), R)
    out["counters"] = dict(zip(COUNTER_NAMES, (int(c) for c in counters)))
    out["checklist"] = checklist[: out["counters"]["checklist_size"]].copy()
    out["construction_cost"], out["final_cost"] = float(costs[0]), float(costs[1])
    return out


def ref_two_opt(matrix, ls, tour, checklist, limit):
    matrix = np.ascontiguousarray(matrix, dtype=np.float64)
    n = matrix.shape[0]
    tour = np.ascontiguousarray(tour, dtype=np.uint32).copy()
    checklist = np.ascontiguousarray(checklist, dtype=np.uint32)
    ls = np.ascontiguousarray(ls, dtype=np.uint32)
    cost, evals = ctypes.c_double(), ctypes.c_uint64()
    R = ref()
    _check(R.ref_two_opt(n, matrix.reshape(-1), ls.reshape(-1), ls.shape[1], tour, checklist, checklist.size,
                         limit, ctypes.byref(cost), ctypes.byref(evals)), R)
    return tour, cost.value, evals.value


def ref_select_next(matrix, primary_row, backup_row, products, visited, current, u):
    matrix = np.ascontiguousarray(matrix, dtype=np.float64)
    chosen, stage = ctypes.c_uint32(), ctypes.c_int()
    primary_row = np.ascontiguousarray(primary_row, dtype=np.uint32)
    backup_row = np.ascontiguousarray(backup_row, dtype=np.uint32)
    R = ref()
    _check(R.ref_select_next(matrix.shape[0], matrix.reshape(-1), primary_row.size, backup_row.size, primary_row,
                             backup_row if backup_row.size else np.zeros(1, np.uint32),
                             np.ascontiguousarray(products, dtype=np.float64),
                             np.ascontiguousarray(visited, dtype=np.uint8), current, u,
                             ctypes.byref(chosen), ctypes.byref(stage)), R)
    return chosen.value, stage.value


def _edge_array(edges):
    e = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1, 2))
    return (e if e.size else np.zeros((1, 2), dtype=np.uint32)), len(edges)


def ref_two_opt_hard(matrix, ls, tour, checklist, limit, edges):
    """The reference's checklist_two_opt inside the graph of `edges` (its
    SparseUndirectedGraph): (tour, cost, counters) with counters processed,
    checks, evaluations, accepted, reactivations, constraint rejections,
    limit reached."""
    matrix = np.ascontiguousarray(matrix, dtype=np.float64)
    n = matrix.shape[0]
    tour = np.ascontiguousarray(tour, dtype=np.uint32).copy()
    checklist = np.ascontiguousarray(checklist, dtype=np.uint32)
    ls = np.ascontiguousarray(ls, dtype=np.uint32)
    e, count = _edge_array(edges)
    cost, counters = ctypes.c_double(), np.zeros(7, dtype=np.uint64)
    R = ref()
    _check(R.ref_two_opt_hard(n, matrix.reshape(-1), ls.reshape(-1), ls.shape[1], tour,
                              checklist if checklist.size else np.zeros(1, np.uint32), checklist.size, limit,
                              e.reshape(-1), count, ctypes.byref(cost), counters), R)
    return tour, cost.value, counters.tolist()


def ref_select_next_hard(matrix, primary_row, backup_row, products, visited, current, u, tour, edges):
    matrix = np.ascontiguousarray(matrix, dtype=np.float64)
    chosen, stage = ctypes.c_uint32(), ctypes.c_int()
    primary_row = np.ascontiguousarray(primary_row, dtype=np.uint32)
    backup_row = np.ascontiguousarray(backup_row, dtype=np.uint32)
    e, count = _edge_array(edges)
    R = ref()
    _check(R.ref_select_next_hard(matrix.shape[0], matrix.reshape(-1), primary_row.size, backup_row.size,
                                  primary_row, backup_row if backup_row.size else np.zeros(1, np.uint32),
                                  np.ascontiguousarray(products, dtype=np.float64),
                                  np.ascontiguousarray(visited, dtype=np.uint8), current, u,
                                  np.ascontiguousarray(tour, dtype=np.uint32), e.reshape(-1), count,
                                  ctypes.byref(chosen), ctypes.byref(stage)), R)
    return chosen.value, stage.value


def ref_trail_limits(width, p_best, retention, cost):
    lo, hi = ctypes.c_double(), ctypes.c_double()
    R = ref()
    _check(R.ref_trail_limits(width, p_best, retention, cost, ctypes.byref(lo), ctypes.byref(hi)), R)
    return lo.value, hi.value


def ref_pheromone_run(matrix, primary, initial, parents, retention, minima, maxima, increments, beta=1.0):
    matrix = np.ascontiguousarray(matrix, dtype=np.float64)
    n = matrix.shape[0]
    primary = np.ascontiguousarray(primary, dtype=np.uint32)
    pw = primary.shape[1]
    parents = np.ascontiguousarray(parents, dtype=np.uint32)
    steps = parents.shape[0]
    trails = np.zeros((steps, n * pw))
    defaults = np.zeros(steps)
    products = np.zeros((steps, n * pw))
    R = ref()
    _check(R.ref_pheromone_run(n, matrix.reshape(-1), primary.reshape(-1), pw, initial, steps, parents.reshape(-1),
                               retention, np.ascontiguousarray(minima, dtype=np.float64),
                               np.ascontiguousarray(maxima, dtype=np.float64),
                               np.ascontiguousarray(increments, dtype=np.float64), beta, trails.reshape(-1),
                               defaults, products.reshape(-1)), R)
    return trails, defaults, products


def ref_score_cpu(opcode, operand, constants, features, masks, feature_spec_id=1):
    opcode = np.ascontiguousarray(opcode, dtype=np.uint8)
    operand = np.ascontiguousarray(operand, dtype=np.uint8)
    constants = np.ascontiguousarray(constants, dtype=np.float32)
    features = np.ascontiguousarray(features, dtype=np.float32)
    masks = np.ascontiguousarray(masks, dtype=np.uint32)
    colonies = masks.size
    scores = np.zeros(colonies * 32, dtype=np.float32)
    actions = (ctypes.c_int32 * colonies)()
    R = ref()
    _check(R.ref_score_cpu(opcode, operand, opcode.size, constants if constants.size else np.zeros(1, np.float32),
                           constants.size, feature_spec_id, features.reshape(-1), masks, colonies, scores,
                           actions), R)
    return scores.reshape(colonies, 32), np.array(actions[:], dtype=np.int32)



# ----------------------------------------------------------------------------
# Control layer.

def tour_ident(tour):
    tour = np.ascontiguousarray(tour, dtype=np.uint32)
    out = np.zeros(2, dtype=np.uint64)
    lib().oracle_tour_ident(tour.size, tour, out)
    return int(out[0]), int(out[1])


def canonical_compare(a, b) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return int(lib().oracle_canonical_compare(a.size, a, np.ascontiguousarray(b, dtype=np.uint32)))


def sample_nodes(n, instance_key):
    out = np.zeros(64, dtype=np.uint32)
    lib().oracle_sample_nodes(n, instance_key, out)
    return out[: min(n, 64)]


def remaining_node(rank, existing) -> int:
    existing = np.ascontiguousarray(existing, dtype=np.uint32)
    return int(lib().oracle_remaining_node(rank, existing if existing.size else np.zeros(1, np.uint32),
                                           existing.size))


def build_region(tour, primary, key, batch, mode, region):
    tour = np.ascontiguousarray(tour, dtype=np.uint32)
    primary = np.ascontiguousarray(primary, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    lib().oracle_build_region(tour.size, tour, primary.reshape(-1), primary.shape[1], key, batch, mode, region, out)
    return out[: min(tour.size, 16)]


def baseline_action(policy: Baseline, feedback: Feedback, batch, uniform, mask) -> int:
    return int(lib().oracle_baseline_action(ctypes.byref(policy), ctypes.byref(feedback), batch, uniform, mask))


def baseline_uniform(policy: Baseline, key, batch) -> float:
    return float(lib().oracle_baseline_uniform(ctypes.byref(policy), key, batch))


def factorial_mask(variant, policy: Baseline, feedback: Feedback, batch, uniform, legal) -> int:
    return int(lib().oracle_factorial_mask(variant, ctypes.byref(policy), ctypes.byref(feedback), batch, uniform,
                                           legal))


def update_feedback(feedback: Feedback, improved, returned, work, ants) -> None:
    lib().oracle_update_feedback(ctypes.byref(feedback), int(bool(improved)), returned, work, ants)


def restart_feedback(feedback: Feedback) -> None:
    lib().oracle_restart_feedback(ctypes.byref(feedback))


def score_controller(controller: Controller, features, mask):
    features = np.ascontiguousarray(features, dtype=np.float32).reshape(12 * 32)
    scores = np.zeros(32, dtype=np.float32)
    action = ctypes.c_uint32()
    L = lib()
    _check(L.oracle_score_controller(ctypes.addressof(controller), features, mask, scores, ctypes.byref(action)), L)
    return scores, action.value


def archive_update(xy, global_tour, entries, completed):
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    entries = np.ascontiguousarray(entries, dtype=np.uint32).reshape(-1, n)
    completed = np.ascontiguousarray(completed, dtype=np.uint32).reshape(-1, n)
    out = np.zeros((4, n), dtype=np.uint32)
    count = ctypes.c_uint32()
    L = lib()
    _check(L.oracle_archive_update(n, xy, np.ascontiguousarray(global_tour, dtype=np.uint32), entries.shape[0],
                                   entries.reshape(-1) if entries.size else np.zeros(1, np.uint32),
                                   completed.shape[0], completed.reshape(-1) if completed.size else np.zeros(1, np.uint32),
                                   out.reshape(-1), ctypes.byref(count)), L)
    return out[: count.value]


def action_snapshot(xy, primary, local_scale, epsilon, samples, active, archive, trails, default_trail, minimum,
                    maximum, feedback: Feedback, elapsed, key, batch, mask=0xFFFFFFFF):
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    primary = np.ascontiguousarray(primary, dtype=np.uint32)
    archive = np.ascontiguousarray(archive, dtype=np.uint32).reshape(-1, n)
    samples = np.ascontiguousarray(samples, dtype=np.uint32)
    features = np.zeros(12 * 32, dtype=np.float32)
    regions = np.zeros(128, dtype=np.uint32)
    alt, legal = ctypes.c_uint32(), ctypes.c_uint32()
    L = lib()
    _check(L.oracle_action_snapshot(n, xy, primary.shape[1], primary.reshape(-1),
                                    np.ascontiguousarray(local_scale, dtype=np.float64), epsilon, samples,
                                    np.ascontiguousarray(active, dtype=np.uint32), archive.shape[0],
                                    archive.reshape(-1), np.ascontiguousarray(trails, dtype=np.float64).reshape(-1),
                                    default_trail, minimum, maximum, ctypes.byref(feedback), elapsed, key, batch,
                                    mask, features, regions, ctypes.byref(alt), ctypes.byref(legal)), L)
    return {"features": features.reshape(12, 32), "regions": regions.reshape(2, 4, 16), "alternative": alt.value,
            "mask": legal.value}


CONTROL_MODES = {"program": 1, "controller": 2, "baseline": 3, "factorial": 4}
FACTORIAL_VARIANTS = {"M00": 0, "M10": 1, "M01": 2, "M11": 3}


def run_controlled(xy_instances, instance_keys, seeds, settings, evaluation_limit, mode, programs=None,
                   controllers=None, baseline=None, variant="M11", experiment_mask=0xFFFFFFFF, threads=0,
                   trace=False, force_collisions=False, or_opt=False, stop_batch=0, fork=None, graphs=None,
                   multipliers=None, progress=None, rows=None, escape=None):
    """Controlled FACO: colony c uses instance c % instances and decider c // instances.

    multipliers [instances, n, pw] (float32): learned heuristic multipliers,
    as for run_faco. progress [batches] (float64): the progress feature of
    each batch (a time-budget run's elapsed share) instead of completed FE
    over evaluation_limit. stop_batch (0 = all) ends the run after that many batches, the progress
    feature still over evaluation_limit; fork = (batch, seed, action) forces
    that batch's action and re-keys every colony from its ants on. `escape`
    as for run_faco."""
    xy_instances = np.ascontiguousarray(xy_instances, dtype=np.int32)
    instances, n = xy_instances.shape[0], xy_instances.shape[1]
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    colonies = seeds.size
    keys = np.ascontiguousarray(instance_keys, dtype=np.uint64)
    tours = np.zeros((colonies, n), dtype=np.uint32)
    costs = np.zeros(colonies, dtype=np.int64)
    work = np.zeros((colonies, 2), dtype=np.uint64)
    rejections = np.zeros(colonies, dtype=np.uint64)
    escape_counters = np.zeros((colonies, len(ESCAPE_COUNTERS)), dtype=np.uint64)
    states = (ControlOut * colonies)()
    elapsed = ctypes.c_double()
    norm = prepare(xy_instances[0], settings)["settings"]
    batches = evaluation_limit // norm.ants
    if stop_batch:
        batches = min(batches, stop_batch)
    traces = {}
    if trace:
        traces = {"features": np.zeros((batches, 12, colonies, 32), dtype=np.float32),
                  "scores": np.zeros((batches, colonies, 32), dtype=np.float32),
                  "masks": np.zeros((batches, colonies), dtype=np.uint32),
                  "actions": np.zeros((batches, colonies), dtype=np.int32),
                  "alternatives": np.zeros((batches, colonies), dtype=np.uint32),
                  "regions": np.zeros((batches, colonies, 2, 4, 16), dtype=np.uint32),
                  "costs": np.zeros((batches, colonies, 3), dtype=np.int64)}
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None  # noqa: E731
    prog_arr = ctrl_arr = None
    deciders = 1
    if programs is not None:
        prog_arr = (Program * len(programs))(*programs)
        deciders = len(programs)
    if controllers is not None:
        ctrl_arr = (Controller * len(controllers))(*controllers)
        deciders = len(controllers)
    policy = baseline if baseline is not None else Baseline()
    g_off, g_nb, g_stride = _graph_arrays(graphs, instances)   # kept alive through the call
    prog = None
    if progress is not None:
        prog = np.ascontiguousarray(progress, dtype=np.float64)
        if prog.size < batches:
            raise ValueError("one progress value per batch")
    mult = None
    if multipliers is not None:
        mult = np.ascontiguousarray(multipliers, dtype=np.float32)
        if mult.shape[0] != instances:
            raise ValueError("one multiplier table per instance")
    spec_rows = _spec_rows(rows, instances)
    L = lib()
    _check(L.oracle_run_controlled(
        n, instances, xy_instances.reshape(-1), keys, colonies, seeds, ctypes.byref(settings), evaluation_limit,
        CONTROL_MODES[mode], deciders, ctypes.addressof(prog_arr) if prog_arr is not None else None,
        ctypes.addressof(ctrl_arr) if ctrl_arr is not None else None, ctypes.byref(policy),
        FACTORIAL_VARIANTS[variant], experiment_mask, int(force_collisions), threads, tours.reshape(-1), costs,
        work.reshape(-1),
        ctypes.addressof(states), ptr(traces.get("features")), ptr(traces.get("scores")),
        ptr(traces.get("masks")), ptr(traces.get("actions")), ptr(traces.get("alternatives")),
        ptr(traces.get("regions")), ptr(traces.get("costs")), ctypes.byref(elapsed), int(bool(or_opt)),
        stop_batch, fork[0] if fork else 0xFFFFFFFF, fork[1] if fork else 0, fork[2] if fork else -1,
        ptr(g_off), ptr(g_nb), g_stride, ptr(mult), ptr(prog), ptr(rejections), ptr(spec_rows),
        ESCAPE_MODES[escape], ptr(escape_counters)), L)
    out = {"tours": tours, "costs": costs, "work": work, "search_seconds": elapsed.value,
           "constraint_rejections": rejections, "escape_counters": escape_counters,
           "states": [{"archive_size": s.archive_size, "stagnant_batches": s.stagnant_batches,
                       "epoch_batches": s.epoch_batches, "restarts": s.restarts, "return_rate": s.return_rate,
                       "ls_work": s.ls_work} for s in states]}
    out.update({"trace_" + k: v for k, v in traces.items()})
    return out


def _colony_arrays(n, pw, costs, tours, trails):
    costs = np.array(costs, dtype=np.int64)
    tours = np.ascontiguousarray(np.asarray(tours, dtype=np.uint32).reshape(3, n)).copy()
    trails = np.ascontiguousarray(np.asarray(trails, dtype=np.float64).reshape(n * pw)).copy()
    return costs, tours, trails


def colony_select(xy, settings, key, batch, ant_tours, ant_costs, costs, tours, trails, default_trail):
    """One batch's reduce/select and pheromone update on explicit colony
    state. costs = (epoch, global, parent), tours = [parent, epoch, global]
    ([3][n]), trails [n][pw]. Returns the state after the batch with the
    bounds, products, iteration best, source coin and its uniform."""
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    norm = prepare(xy, settings)["settings"]
    ant_tours = np.ascontiguousarray(ant_tours, dtype=np.uint32)
    ant_costs = np.ascontiguousarray(ant_costs, dtype=np.int64)
    costs, tours, trails = _colony_arrays(n, norm.primary_width, costs, tours, trails)
    default = np.array([default_trail], dtype=np.float64)
    bounds = np.zeros(2, dtype=np.float64)
    products = np.zeros(n * norm.primary_width, dtype=np.float64)
    best, epoch_source, uniform = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
    L = lib()
    _check(L.oracle_colony_select(n, xy, ctypes.byref(settings), key, batch, ant_costs.size, ant_tours.reshape(-1),
                                  ant_costs, costs, tours, trails, default, bounds, products, ctypes.byref(best),
                                  ctypes.byref(epoch_source), ctypes.byref(uniform)), L)
    return {"costs": costs, "parent": tours[0], "epoch": tours[1], "global": tours[2],
            "trails": trails.reshape(n, norm.primary_width), "default_trail": float(default[0]),
            "minimum": float(bounds[0]), "maximum": float(bounds[1]),
            "products": products.reshape(n, norm.primary_width), "iteration_best": best.value,
            "source_is_epoch": bool(epoch_source.value), "source_uniform": uniform.value}


def control_restart(xy, settings, archive, archive_costs, alternative, initializing, costs, tours, trails,
                    default_trail, feedback: Feedback):
    """One restart transaction on explicit colony state (same layout as
    colony_select) with an archive [count][n]; `feedback` is updated in place."""
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    norm = prepare(xy, settings)["settings"]
    archive = np.ascontiguousarray(archive, dtype=np.uint32)
    archive_costs = np.ascontiguousarray(archive_costs, dtype=np.int64)
    costs, tours, trails = _colony_arrays(n, norm.primary_width, costs, tours, trails)
    default = np.array([default_trail], dtype=np.float64)
    bounds = np.zeros(2, dtype=np.float64)
    products = np.zeros(n * norm.primary_width, dtype=np.float64)
    identities = np.zeros(4, dtype=np.uint64)
    L = lib()
    _check(L.oracle_control_restart(n, xy, ctypes.byref(settings), archive_costs.size, archive.reshape(-1),
                                    archive_costs, alternative, int(bool(initializing)), costs, tours, trails,
                                    default, bounds, products, ctypes.byref(feedback), identities), L)
    return {"costs": costs, "parent": tours[0], "epoch": tours[1], "global": tours[2],
            "trails": trails.reshape(n, norm.primary_width), "default_trail": float(default[0]),
            "minimum": float(bounds[0]), "maximum": float(bounds[1]),
            "products": products.reshape(n, norm.primary_width),
            "active_identity": (int(identities[0]), int(identities[1])),
            "epoch_identity": (int(identities[2]), int(identities[3]))}


# Reference checker, control layer.

def _ref_control():
    R = ref()
    if not getattr(R, "_control_bound", False):
        R.ref_fingerprint.argtypes = [ctypes.c_uint32, u32p, u64p]
        R.ref_canonical_compare.restype = ctypes.c_int
        R.ref_canonical_compare.argtypes = [ctypes.c_uint32, u32p, u32p]
        R.ref_sample_nodes.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p]
        R.ref_build_region.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_uint64,
                                       ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p]
        R.ref_validate_baseline.argtypes = [ctypes.POINTER(Baseline), ctypes.c_uint32]
        R.ref_baseline_action.restype = ctypes.c_uint32
        R.ref_baseline_action.argtypes = [ctypes.POINTER(Baseline), ctypes.POINTER(Feedback), ctypes.c_uint32,
                                          ctypes.c_double, ctypes.c_uint32]
        R.ref_factorial_mask.restype = ctypes.c_uint32
        R.ref_factorial_mask.argtypes = [ctypes.c_uint32, ctypes.POINTER(Baseline), ctypes.POINTER(Feedback),
                                         ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32]
        R.ref_update_feedback.argtypes = [ctypes.POINTER(Feedback), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_uint32]
        R.ref_restart_feedback.argtypes = [ctypes.POINTER(Feedback)]
        R.ref_archive_update.argtypes = [ctypes.c_uint32, f64p, u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p,
                                         u32p, ctypes.POINTER(ctypes.c_uint32)]
        R.ref_action_snapshot.argtypes = [ctypes.c_uint32, f64p, ctypes.c_uint32, u32p, f64p, ctypes.c_double, u32p,
                                          ctypes.c_uint32, u32p, ctypes.c_uint32, u32p, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_uint32, u32p, f64p,
                                          ctypes.c_double, ctypes.c_double, ctypes.POINTER(Feedback),
                                          ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, f32p,
                                          u32p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                          f64p, ctypes.POINTER(ctypes.c_double), u32p,
                                          ctypes.POINTER(ctypes.c_uint32)]
        R._control_bound = True
    return R


def ref_fingerprint(tour):
    tour = np.ascontiguousarray(tour, dtype=np.uint32)
    out = np.zeros(2, dtype=np.uint64)
    _ref_control().ref_fingerprint(tour.size, tour, out)
    return int(out[0]), int(out[1])


def ref_canonical_compare(a, b) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return int(_ref_control().ref_canonical_compare(a.size, a, np.ascontiguousarray(b, dtype=np.uint32)))


def ref_sample_nodes(n, key):
    out = np.zeros(64, dtype=np.uint32)
    _ref_control().ref_sample_nodes(n, key, out)
    return out[: min(n, 64)]


def ref_build_region(tour, primary, key, batch, mode, region):
    tour = np.ascontiguousarray(tour, dtype=np.uint32)
    primary = np.ascontiguousarray(primary, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    _ref_control().ref_build_region(tour.size, tour, primary.reshape(-1), primary.shape[1], key, batch, mode,
                                    region, out)
    return out[: min(tour.size, 16)]


def ref_validate_baseline(policy: Baseline, mask=0xFFFFFFFF) -> bool:
    return _ref_control().ref_validate_baseline(ctypes.byref(policy), mask) == 0


def ref_baseline_action(policy, feedback, batch, uniform, mask) -> int:
    return int(_ref_control().ref_baseline_action(ctypes.byref(policy), ctypes.byref(feedback), batch, uniform,
                                                  mask))


def ref_factorial_mask(variant, policy, feedback, batch, uniform, legal) -> int:
    return int(_ref_control().ref_factorial_mask(variant, ctypes.byref(policy), ctypes.byref(feedback), batch,
                                                 uniform, legal))


def ref_update_feedback(feedback, improved, returned, work, ants) -> None:
    _ref_control().ref_update_feedback(ctypes.byref(feedback), int(bool(improved)), returned, work, ants)


def ref_restart_feedback(feedback) -> None:
    _ref_control().ref_restart_feedback(ctypes.byref(feedback))


def ref_archive_update(matrix, global_tour, entries, completed):
    matrix = np.ascontiguousarray(matrix, dtype=np.float64)
    n = matrix.shape[0]
    entries = np.ascontiguousarray(entries, dtype=np.uint32).reshape(-1, n)
    completed = np.ascontiguousarray(completed, dtype=np.uint32).reshape(-1, n)
    out = np.zeros((4, n), dtype=np.uint32)
    count = ctypes.c_uint32()
    R = _ref_control()
    _check(R.ref_archive_update(n, matrix.reshape(-1), np.ascontiguousarray(global_tour, dtype=np.uint32),
                                entries.shape[0], entries.reshape(-1) if entries.size else np.zeros(1, np.uint32),
                                completed.shape[0],
                                completed.reshape(-1) if completed.size else np.zeros(1, np.uint32),
                                out.reshape(-1), ctypes.byref(count)), R)
    return out[: count.value]


def ref_action_snapshot(matrix, primary, local_scale, epsilon, samples, active, archive_tours, pheromone_initial,
                        retention, evaporation_floor, deposit_parents, deposit_amounts, minimum, maximum,
                        feedback, elapsed, key, batch, mask=0xFFFFFFFF):
    matrix = np.ascontiguousarray(matrix, dtype=np.float64)
    n = matrix.shape[0]
    primary = np.ascontiguousarray(primary, dtype=np.uint32)
    pw = primary.shape[1]
    archive_tours = np.ascontiguousarray(archive_tours, dtype=np.uint32).reshape(-1, n)
    deposit_parents = np.ascontiguousarray(deposit_parents, dtype=np.uint32).reshape(-1, n)
    samples = np.ascontiguousarray(samples, dtype=np.uint32)
    features = np.zeros(12 * 32, dtype=np.float32)
    regions = np.zeros(128, dtype=np.uint32)
    alt, legal, count = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    trails = np.zeros(n * pw)
    default = ctypes.c_double()
    archive_out = np.zeros((4, n), dtype=np.uint32)
    R = _ref_control()
    _check(R.ref_action_snapshot(n, matrix.reshape(-1), pw, primary.reshape(-1),
                                 np.ascontiguousarray(local_scale, dtype=np.float64), epsilon, samples, samples.size,
                                 np.ascontiguousarray(active, dtype=np.uint32), archive_tours.shape[0],
                                 archive_tours.reshape(-1), pheromone_initial, retention, evaporation_floor,
                                 deposit_parents.shape[0],
                                 deposit_parents.reshape(-1) if deposit_parents.size else np.zeros(1, np.uint32),
                                 np.ascontiguousarray(deposit_amounts, dtype=np.float64) if len(deposit_amounts)
                                 else np.zeros(1), minimum, maximum, ctypes.byref(feedback), elapsed, key, batch,
                                 mask, features, regions, ctypes.byref(alt), ctypes.byref(legal), trails,
                                 ctypes.byref(default), archive_out.reshape(-1), ctypes.byref(count)), R)
    return {"features": features.reshape(12, 32), "regions": regions.reshape(2, 4, 16), "alternative": alt.value,
            "mask": legal.value, "trails": trails.reshape(n, pw), "default": default.value,
            "archive": archive_out[: count.value]}
