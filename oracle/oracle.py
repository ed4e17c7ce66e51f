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
  the parts above in tests/test_oracle_colony.py.
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
                                          ctypes.c_uint32, ctypes.c_uint64, u32p, u32p, u32p, u32p, i64p, u64p]
        L.oracle_select_next.argtypes = [ctypes.c_uint32, i32p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p,
                                         f64p, u8p, ctypes.c_uint32, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int)]
        L.oracle_run_faco.argtypes = [ctypes.c_uint32, ctypes.c_uint32, i32p, u64p, ctypes.c_uint32, u64p,
                                      ctypes.POINTER(Settings), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                      u32p, i64p, u64p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_double)]
        L.oracle_max_threads.restype = ctypes.c_int
        L.oracle_two_opt.argtypes = [ctypes.c_uint32, i32p, u32p, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32,
                                     ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint64)]
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


def construct_ls(xy, ls, parent, visit, mne_target, limit) -> dict:
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
                                 counters), L)
    out["counters"] = dict(zip(COUNTER_NAMES, (int(c) for c in counters)))
    out["checklist"] = checklist[: out["counters"]["checklist_size"]].copy()
    out["construction_cost"], out["final_cost"] = int(costs[0]), int(costs[1])
    return out


def two_opt(xy, ls, tour, checklist, limit):
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1)
    n = xy.size // 2
    ls = np.ascontiguousarray(ls, dtype=np.uint32)
    tour = np.ascontiguousarray(tour, dtype=np.uint32).copy()
    checklist = np.ascontiguousarray(checklist, dtype=np.uint32)
    cost, evals = ctypes.c_int64(), ctypes.c_uint64()
    L = lib()
    _check(L.oracle_two_opt(n, xy, ls.reshape(-1), ls.shape[1], tour, checklist, checklist.size, limit,
                            ctypes.byref(cost), ctypes.byref(evals)), L)
    return tour, cost.value, evals.value


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


def run_faco(xy_instances, instance_keys, seeds, settings, batches, mne_target=8, threads=0, trace=False):
    """Fixed-iteration FACO over colonies (colony i uses instance i % instances)."""
    xy_instances = np.ascontiguousarray(xy_instances, dtype=np.int32)
    instances, n = xy_instances.shape[0], xy_instances.shape[1]
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    colonies = seeds.size
    keys = np.ascontiguousarray(instance_keys, dtype=np.uint64)
    tours = np.zeros((colonies, n), dtype=np.uint32)
    costs = np.zeros(colonies, dtype=np.int64)
    work = np.zeros((colonies, 2), dtype=np.uint64)
    elapsed = ctypes.c_double()
    trace_parent = trace_trails = trace_costs = None
    if trace:
        norm = prepare(xy_instances[0], settings)["settings"]
        trace_parent = np.zeros((batches, n), dtype=np.uint32)
        trace_trails = np.zeros((batches, n * norm.primary_width), dtype=np.float64)
        trace_costs = np.zeros((batches, 3), dtype=np.int64)
    L = lib()
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None  # noqa: E731
    _check(L.oracle_run_faco(n, instances, xy_instances.reshape(-1), keys, colonies, seeds,
                             ctypes.byref(settings), batches, mne_target, threads, tours.reshape(-1), costs,
                             work.reshape(-1), ptr(trace_parent), ptr(trace_trails), ptr(trace_costs),
                             ctypes.byref(elapsed)), L)
    out = {"tours": tours, "costs": costs, "work": work, "search_seconds": elapsed.value}
    if trace:
        out.update(trace_parent=trace_parent, trace_trails=trace_trails, trace_costs=trace_costs)
    return out


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

