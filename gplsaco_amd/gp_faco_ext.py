"""Drop-in Python surface of the reference's native module ``gp_faco_ext``.

Names, keyword-only flags, dtype strictness and result-dict keys follow the
reference pybind11 module (reference cpp/src/bindings.cpp:500-1046); every
call goes through the C ABI of libfaco_hip.so (include/faco_engine.h). A
caller switches with ``from gplsaco_amd import gp_faco_ext``.

Differences from the reference, by design of this engine:
- distances are TSPLIB EUC_2D on integer coordinates in [0, 32767] (the
  reference uses continuous FP64); ``numeric_backend`` says so;
- ``score_cuda`` runs the HIP kernel (the name is kept for callers).
"""
from __future__ import annotations

import ctypes
import json
import struct
import time

import numpy as np

from . import _native
from ._native import BatchProfile, DeadlineResult, DeadlineSpec
from ._native import (BaselinePolicy, Checkpoints, ControlSpec, ControlState, ControlTrace, Controller, Fork, Program,
                      Result, Settings,
                      TuneEntry, check, lib, ptr)

numeric_backend = "euc2d_int32"
campaign_interface_version = 1

# Escape counters in the reference's EscapeStats order (escape_slots.hpp:70-75)
ESCAPE_COUNTERS = ("construction_opportunities", "construction_gates", "construction_replaced_slots",
                   "escape_relocations", "ls_anchor_nodes", "ls_replaced_slots", "old_view_reactivations",
                   "anchor_reactivations", "new_edges", "construction_capacity_rejections", "ls_capacity_rejections")

_SETTING_FIELDS = ("ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
                   "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")


class FixedFacoSettings:
    """Read/write settings record (reference bindings.cpp:508-519)."""

    def __init__(self):
        self.ants = 0
        self.primary_width = 16
        self.backup_width = 64
        self.ls_width = 20
        self.beta = 1.0
        self.retention = 0.5
        self.p_best = 0.1
        self.epoch_source_probability = 0.01
        self.ls_evaluation_limit = 100000
        self.initial_ls_evaluation_limit = 100000

    def _native(self) -> Settings:
        s = Settings()
        for name in _SETTING_FIELDS:
            value = getattr(self, name)
            if isinstance(value, bool):
                raise TypeError(f"{name} must be a number, not bool")
            setattr(s, name, value)
        return s

    @classmethod
    def _from_native(cls, s: Settings) -> "FixedFacoSettings":
        out = cls()
        for name in _SETTING_FIELDS:
            setattr(out, name, getattr(s, name))
        return out

    def __repr__(self):
        return "FixedFacoSettings(" + ", ".join(f"{k}={getattr(self, k)!r}" for k in _SETTING_FIELDS) + ")"


def _coordinates(coordinates) -> np.ndarray:
    if not isinstance(coordinates, np.ndarray) or coordinates.dtype != np.float64 or coordinates.ndim != 2 \
            or coordinates.shape[1] != 2 or not coordinates.flags.c_contiguous:
        raise TypeError("coordinates must be a C-contiguous float64 array of shape [n, 2]")
    return coordinates


def _keys(values, name) -> np.ndarray:
    if not isinstance(values, np.ndarray) or values.dtype != np.uint64 or values.ndim != 1 \
            or not values.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous uint64 array")
    return values


def _unsigned(value, name, bits=64) -> int:
    if type(value) is not int:
        raise ValueError(f"{name} must be a non-negative integer, not {type(value).__name__}")
    if value < 0 or value >= 1 << bits:
        raise ValueError(f"{name} out of the uint{bits} range")
    return value


# ---------------------------------------------------------------------------
# Program IR (reference bindings.cpp:22-56, python/gp_faco/program_ir.py:20-130)

def read_program(dictionary) -> Program:
    if not isinstance(dictionary, dict):
        raise TypeError("program must be a dict")
    if len(dictionary) != 6 and not (len(dictionary) == 7 and "program_id" in dictionary):
        raise ValueError("IR field count is wrong")
    for name in ("ir_version", "numeric_spec_id", "feature_spec_id"):
        if type(dictionary.get(name)) is not int:
            raise ValueError("IR versions must be integers")
    for name in ("opcode", "operand", "constant_bits"):
        seq = dictionary.get(name)
        if not isinstance(seq, (list, tuple)) or any(type(v) is not int for v in seq):
            raise ValueError("IR opcodes and bit patterns must be integers")
    opcodes, operands, constants = dictionary["opcode"], dictionary["operand"], dictionary["constant_bits"]
    if not opcodes or len(opcodes) > 63 or len(operands) != len(opcodes) or len(constants) > 63:
        raise ValueError("IR array size is wrong")
    if any(not 0 <= v < 256 for v in list(opcodes) + list(operands)) or any(not 0 <= v < 1 << 32 for v in constants):
        raise ValueError("IR values out of range")
    for name in ("ir_version", "numeric_spec_id", "feature_spec_id"):
        if not 0 <= dictionary[name] < 1 << 16:
            raise ValueError("IR versions out of range")
    p = Program()
    p.ir_version = dictionary["ir_version"]
    p.numeric_spec_id = dictionary["numeric_spec_id"]
    p.feature_spec_id = dictionary["feature_spec_id"]
    p.length = len(opcodes)
    p.constants_count = len(constants)
    for i, (op, operand) in enumerate(zip(opcodes, operands)):
        p.opcode[i] = op
        p.operand[i] = operand
    for i, bits in enumerate(constants):
        p.constants[i] = struct.unpack("<f", struct.pack("<I", bits))[0]
    return p


def _score(program, features, masks, entry):
    p = read_program(program)
    if not isinstance(features, np.ndarray) or features.dtype != np.float32 or features.ndim != 3 \
            or features.shape[0] != 12 or features.shape[2] != 32 or not features.flags.c_contiguous:
        raise TypeError("features must be C-contiguous float32 [12, colonies, 32]")
    if not isinstance(masks, np.ndarray) or masks.dtype != np.uint32 or masks.ndim != 1 \
            or masks.shape[0] != features.shape[1] or not masks.flags.c_contiguous:
        raise TypeError("masks must be C-contiguous uint32 [colonies]")
    colonies = masks.shape[0]
    scores = np.zeros((colonies, 32), dtype=np.float32)
    actions = np.zeros(colonies, dtype=np.int32)
    check(getattr(lib(), entry)(ctypes.byref(p), features.reshape(-1), masks, colonies, scores.reshape(-1), actions))
    return {"scores": scores, "actions": actions}


def score_cpu(program, features, masks):
    """Host scorer (reference cpp/src/program_cpu.cpp:79-116)."""
    return _score(program, features, masks, "faco_score_cpu")


def score_cuda(program, features, masks):
    """HIP scorer (reference cuda/gp_score.cu:52-77); bit-identical to score_cpu."""
    return _score(program, features, masks, "faco_score_gpu")


score_hip = score_cuda


# ---------------------------------------------------------------------------

# ---------------------------------------------------------------------------
# Controllers and policies (reference bindings.cpp:56-75,147-220)

def read_controller(dictionary) -> Controller:
    """A program dict (one tree) or {"controller_version": 1, "representation":
    "joint_single" | "conditional_three", "trees": [...]}."""
    if not isinstance(dictionary, dict):
        raise TypeError("controller must be a dict")
    c = Controller()
    if "opcode" in dictionary:
        c.kind = 0
        c.trees[0] = read_program(dictionary)
        return c
    if len(dictionary) != 4 or type(dictionary.get("controller_version")) is not int \
            or dictionary["controller_version"] != 1:
        raise ValueError("unknown controller version or fields")
    kind = dictionary.get("representation")
    if kind not in ("joint_single", "conditional_three"):
        raise ValueError("unknown controller representation")
    c.kind = 1 if kind == "conditional_three" else 0
    trees = dictionary.get("trees")
    if not isinstance(trees, (list, tuple)) or len(trees) != (3 if c.kind else 1):
        raise ValueError("controller tree count is wrong")
    for i, tree in enumerate(trees):
        c.trees[i] = read_program(tree)
    return c


_BASELINE_INT_FIELDS = ("mne_level", "max_mne_level", "region", "restart_period", "stagnation_step",
                        "restart_stagnation", "restart_cooldown")


def read_baseline(dictionary) -> BaselinePolicy:
    if not isinstance(dictionary, dict) or len(dictionary) != 11 or type(dictionary.get("policy_spec_id")) is not int \
            or dictionary["policy_spec_id"] != 1:
        raise ValueError("baseline fields or spec version invalid")
    b = BaselinePolicy()
    kinds = {"static": 0, "rule": 1}
    modes = {"none": 0, "periodic": 1, "bernoulli": 2}
    if dictionary.get("kind") not in kinds:
        raise ValueError("unknown baseline kind")
    if dictionary.get("restart_mode") not in modes:
        raise ValueError("unknown static restart mode")
    b.kind, b.restart_mode = kinds[dictionary["kind"]], modes[dictionary["restart_mode"]]
    for name in _BASELINE_INT_FIELDS:
        setattr(b, name, _unsigned(dictionary.get(name), name, 32))
    probability = dictionary.get("restart_probability")
    if type(probability) not in (int, float):
        raise ValueError("restart probability must be a real number, not bool")
    b.restart_probability = float(probability)
    return b


def baseline_output(b: BaselinePolicy) -> dict:
    return {"policy_spec_id": 1, "kind": "static" if b.kind == 0 else "rule", "mne_level": b.mne_level,
            "max_mne_level": b.max_mne_level, "region": b.region,
            "restart_mode": ("none", "periodic", "bernoulli")[b.restart_mode], "restart_period": b.restart_period,
            "restart_probability": b.restart_probability, "stagnation_step": b.stagnation_step,
            "restart_stagnation": b.restart_stagnation, "restart_cooldown": b.restart_cooldown}


FACTORIAL_VARIANTS = ("M00", "M10", "M01", "M11")


def read_factorial(dictionary):
    if not isinstance(dictionary, dict) or len(dictionary) != 3 or set(dictionary) != {
            "factorial_spec_id", "variant", "baseline_policy"} or type(dictionary["factorial_spec_id"]) is not int \
            or dictionary["factorial_spec_id"] != 1 or not isinstance(dictionary["baseline_policy"], dict):
        raise ValueError("factorial fields or spec version invalid")
    if dictionary["variant"] not in FACTORIAL_VARIANTS:
        raise ValueError("unknown factorial variant")
    return FACTORIAL_VARIANTS.index(dictionary["variant"]), read_baseline(dictionary["baseline_policy"])


def factorial_output(variant: int, baseline: BaselinePolicy) -> dict:
    return {"factorial_spec_id": 1, "variant": FACTORIAL_VARIANTS[variant], "baseline_policy": baseline_output(baseline)}


def _controller_score(controller, features, masks, entry):
    c = read_controller(controller)
    if not isinstance(features, np.ndarray) or features.dtype != np.float32 or features.ndim != 3 \
            or features.shape[0] != 12 or features.shape[2] != 32 or not features.flags.c_contiguous:
        raise TypeError("features must be C-contiguous float32 [12, colonies, 32]")
    if not isinstance(masks, np.ndarray) or masks.dtype != np.uint32 or masks.ndim != 1 \
            or masks.size != features.shape[1] or not masks.flags.c_contiguous:
        raise TypeError("masks must be C-contiguous uint32 [colonies]")
    colonies = masks.size
    scores = np.zeros((colonies, 32), dtype=np.float32)
    actions = np.zeros(colonies, dtype=np.int32)
    check(getattr(lib(), entry)(ctypes.byref(c), features.reshape(-1), masks, colonies, scores.reshape(-1), actions))
    return {"scores": scores, "actions": actions}


def score_controller_cpu(controller, features, masks):
    """Host scorer of a ControllerProgram (the device kernel's twin)."""
    return _controller_score(controller, features, masks, "faco_score_controller_cpu")


def score_controller_cuda(controller, features, masks):
    """score_controller_cuda (reference bindings.cpp:1010-1030) on the HIP device."""
    return _controller_score(controller, features, masks, "faco_score_controller_gpu")


def prepare_common_initial(coordinates, settings: FixedFacoSettings) -> dict:
    xy = _coordinates(coordinates)
    n = xy.shape[0]
    tour = np.zeros(n, dtype=np.uint32)
    cost = ctypes.c_int64()
    cheap, prep = ctypes.c_double(), ctypes.c_double()
    norm = Settings()
    check(lib().faco_prepare_common_initial(xy.reshape(-1), n, ctypes.byref(settings._native()), ctypes.byref(norm),
                                            tour, ctypes.byref(cost), ctypes.byref(cheap), ctypes.byref(prep),
                                            None, None, None, None))
    return {"tour": tour.tolist(), "cost": float(cost.value), "cheap_seconds": cheap.value,
            "preparation_seconds": prep.value}


def prepare_rows(coordinates, settings: FixedFacoSettings) -> dict:
    """Candidate rows and initial tour of the host preparation (as arrays)."""
    xy = _coordinates(coordinates)
    n = xy.shape[0]
    norm = Settings()
    tour = np.zeros(n, dtype=np.uint32)
    cost = ctypes.c_int64()
    cheap, prep = ctypes.c_double(), ctypes.c_double()
    s = settings._native()
    primary = np.zeros(n * max(1, s.primary_width), dtype=np.uint32)
    backup = np.zeros(n * max(1, s.backup_width), dtype=np.uint32)
    ls = np.zeros(n * max(1, s.ls_width), dtype=np.uint32)
    scale = np.zeros(n, dtype=np.float64)
    check(lib().faco_prepare_common_initial(xy.reshape(-1), n, ctypes.byref(s), ctypes.byref(norm), tour,
                                            ctypes.byref(cost), ctypes.byref(cheap), ctypes.byref(prep),
                                            ptr(primary), ptr(backup), ptr(ls), ptr(scale)))
    return {"settings": FixedFacoSettings._from_native(norm),
            "primary": primary[: n * norm.primary_width].reshape(n, norm.primary_width),
            "backup": backup[: n * norm.backup_width].reshape(n, norm.backup_width),
            "ls": ls[: n * norm.ls_width].reshape(n, norm.ls_width),
            "local_scale": scale, "initial_tour": tour, "initial_cost": int(cost.value)}


def cuda_device_info() -> dict:
    device, cus = ctypes.c_int(), ctypes.c_int()
    mem, lds = ctypes.c_size_t(), ctypes.c_size_t()
    name, arch = ctypes.create_string_buffer(256), ctypes.create_string_buffer(64)
    check(lib().faco_device_info(ctypes.byref(device), name, 256, arch, 64, ctypes.byref(cus), ctypes.byref(mem),
                                 ctypes.byref(lds)))
    return {"device": device.value, "name": name.value.decode(), "arch": arch.value.decode(),
            "compute_units": cus.value, "total_memory": mem.value, "max_lds_per_block": lds.value}


class _CheckpointRequest:
    """checkpoint_iterations of the count entry points (reference
    cuda/batch_engine.cu:635-653, bindings.cpp:308-321): the global-best tour
    of every colony after each listed number of completed batches. The engine
    checks order and range; a row of a smaller mixed-size colony ends at its
    own n."""

    def __init__(self, points, colonies, dimension):
        points = list(points)
        if any(type(v) is not int or v < 0 or v >= 1 << 64 for v in points):
            raise ValueError("checkpoint iterations must be non-negative integers")
        self.points = points
        self.iterations = np.ascontiguousarray(points, dtype=np.uint64)
        self.tours = np.zeros((len(points), colonies, dimension), dtype=np.uint32)
        self.struct = Checkpoints()
        self.struct.iterations = self.iterations.ctypes.data_as(ctypes.c_void_p)
        self.struct.count = len(points)
        self.struct.tours = self.tours.ctypes.data_as(ctypes.c_void_p)

    def argument(self):
        return ctypes.byref(self.struct) if self.points else None

    def output(self) -> list:
        out = []
        for k, iterations in enumerate(self.points):
            tours = []
            for row in self.tours[k]:
                end = np.flatnonzero(row == 0xFFFFFFFF)
                tours.append(row[: end[0] if end.size else row.size].tolist())
            out.append({"iterations": iterations, "tours": tours})
        return out


def _production_stages(km, controlled: bool) -> dict:
    """production_kernel_milliseconds under the reference's stage names
    (cpp/include/gp_faco/profiling.hpp:11-19). Stages that run fused in one
    kernel here report under the first name of their group and 0.0 under the
    others: decision = features_and_regions (+ gp_scoring, action_and_restart),
    colony update = reduction (+ pheromone), archive update =
    archive_and_feedback (+ ant_fingerprints)."""
    return {"initialization": km[4] if controlled else 0.0,
            "features_and_regions": km[2] if controlled else 0.0, "gp_scoring": 0.0, "action_and_restart": 0.0,
            "construction_and_ls": km[0], "reduction": km[1], "pheromone": 0.0, "ant_fingerprints": 0.0,
            "archive_and_feedback": km[3] if controlled else 0.0}


def _batch_output(result: Result, tours: np.ndarray, costs: np.ndarray, mode: str, elapsed: float) -> dict:
    items = [{"has_incumbent": True, "tour": tours[i].tolist(), "cost": float(costs[i]),
              "completed_seconds": elapsed} for i in range(tours.shape[0])]
    return {
        "items": items,
        "budget_seconds": None,
        "budget_kind": "search_tour_evaluations",
        "evaluation_limit_per_colony": result.evaluation_limit_per_colony,
        "completed_tour_evaluations_per_colony": result.completed_tour_evaluations_per_colony,
        "total_tour_evaluations": result.total_tour_evaluations,
        "elapsed_seconds": result.actual_seconds,
        "actual_seconds": result.actual_seconds,
        "charged_seconds": 0.0,
        "last_batch_completed_seconds": elapsed,
        "overrun_seconds": 0.0,
        "launched_batches": result.launched_batches,
        "completed_batches": result.completed_batches,
        "discarded_batches": 0,
        "completed_construction_steps": result.completed_construction_steps,
        "completed_ls_evaluations": result.completed_ls_evaluations,
        "allocated_device_bytes": result.allocated_device_bytes,
        "preparation_completed": True,
        "preparation_mode": mode,
        "search_seconds": result.search_seconds,
        "search_gpu_milliseconds": result.search_gpu_milliseconds,
        "persistent_launches": result.persistent_launches,
    }


def _control_spec(mode, programs, controllers, baseline, variant, experiment_mask, force_collisions):
    """The faco_control_spec of a controlled run and the decider arrays it
    points to (keep them alive through the call)."""
    spec = ControlSpec()
    spec.mode = {"program": 1, "controller": 2, "baseline": 3, "factorial": 4}[mode]
    spec.experiment_mask = _unsigned(experiment_mask, "experiment_mask", 32)
    spec.factorial_variant = variant
    spec.force_fingerprint_collisions = 1 if force_collisions else 0
    prog_arr = ctrl_arr = None
    if mode in ("program", "factorial"):
        prog_arr = (Program * len(programs))(*programs)
        spec.programs = ctypes.addressof(prog_arr)
        spec.deciders = len(programs)
    if mode == "controller":
        ctrl_arr = (Controller * len(controllers))(*controllers)
        spec.controllers = ctypes.addressof(ctrl_arr)
        spec.deciders = len(controllers)
    if mode == "baseline":
        spec.deciders = 1
    if baseline is not None:
        spec.baseline = baseline
    return spec, prog_arr, ctrl_arr


def _control_state_dict(st: ControlState) -> dict:
    """One entry of control_states (reference bindings.cpp:394-407)."""
    return {"archive_size": st.archive_size, "stagnant_batches": st.stagnant_batches, "return_rate": st.return_rate,
            "ls_work": st.ls_work, "epoch_batches": st.epoch_batches, "restarts": st.restarts}


class CountedState:
    """Snapshot of a controlled run after a whole number of batches
    (reference cpp/include/gp_faco/counted_state.hpp:1-48, bindings.cpp:541-556).

    The bytes are the reference's container (magic "GFACOST2", 64-bit
    little-endian fields, named buffers); the buffers carry this library's
    device state, so a snapshot resumes only in an engine of this library."""

    def __init__(self, handle):
        self._handle = handle
        self._colonies = None

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            lib().faco_state_destroy(handle)
            self._handle = None

    @staticmethod
    def from_bytes(data) -> "CountedState":
        if not isinstance(data, (bytes, bytearray)):
            raise TypeError("snapshot bytes must be bytes")
        handle = ctypes.c_void_p()
        check(lib().faco_state_from_bytes(bytes(data), len(data), ctypes.byref(handle)))
        return CountedState(handle)

    def to_bytes(self) -> bytes:
        size = ctypes.c_size_t()
        check(lib().faco_state_to_bytes(self._handle, None, 0, ctypes.byref(size)))
        buffer = ctypes.create_string_buffer(max(size.value, 1))
        check(lib().faco_state_to_bytes(self._handle, buffer, size.value, ctypes.byref(size)))
        return buffer.raw[:size.value]

    def describe(self) -> dict:
        size = ctypes.c_size_t()
        check(lib().faco_state_describe(self._handle, None, 0, ctypes.byref(size)))
        buffer = ctypes.create_string_buffer(size.value)
        check(lib().faco_state_describe(self._handle, buffer, size.value, ctypes.byref(size)))
        return json.loads(buffer.value.decode())

    @property
    def colonies(self) -> int:
        if self._colonies is None:
            self._colonies = self.describe()["colonies"]
        return self._colonies

    def select_colony(self, colony) -> "CountedState":
        index = _unsigned(colony, "colony", 32)
        handle = ctypes.c_void_p()
        check(lib().faco_state_select_colony(self._handle, index, ctypes.byref(handle)))
        return CountedState(handle)


class FacoBatchEngine:
    """Fixed-shape multi-colony engine (reference batch_engine.hpp:68-145)."""

    def __init__(self, dimension, colonies, settings, constraint_mode="unrestricted", population_size=1):
        if type(dimension) is not int or type(colonies) is not int:
            raise ValueError("dimension/colonies must be integers")
        population = _unsigned(population_size, "population_size", 32)
        if constraint_mode not in ("unrestricted", "hard", "escape"):
            raise ValueError("unknown constraint mode")
        if not isinstance(settings, FixedFacoSettings):
            raise TypeError("settings must be FixedFacoSettings")
        handle = ctypes.c_void_p()
        check(lib().faco_engine_create(dimension, colonies, ctypes.byref(settings._native()),
                                       {"unrestricted": 0, "hard": 1, "escape": 2}[constraint_mode], population,
                                       ctypes.byref(handle)))
        self.constraint_mode = constraint_mode
        # diagnostic (the reference's BatchDiagnosticControls::disable_escape):
        # an Escape engine's evaluations keep every escape gate shut, which
        # must reproduce the Hard run
        self.disable_escape = False
        self._handle = handle
        self.dimension, self.colonies, self.population_size = dimension, colonies, population
        s = Settings()
        check(lib().faco_engine_settings(self._handle, ctypes.byref(s)))
        self.settings = FixedFacoSettings._from_native(s)
        self._charges = {}
        self._graph_edges = {}   # Hard mode: undirected edge count of each registered graph

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            lib().faco_engine_destroy(handle)
            self._handle = None

    def launch_shape(self, colonies: int = 0) -> dict:
        vals = [ctypes.c_uint32() for _ in range(5)]
        check(lib().faco_engine_launch_shape(self._handle, colonies, *[ctypes.byref(v) for v in vals]))
        w, lds_rows, lds, bpc, per_cu = (v.value for v in vals)
        return {"waves_per_block": w, "neighbour_rows_in_lds": bool(lds_rows), "lds_bytes": lds,
                "blocks_per_colony": bpc, "blocks_per_cu": per_cu}

    def set_local_search(self, mode: str) -> None:
        """'2opt' (the reference's checklist 2-opt, default) or '2opt+oropt'."""
        check(lib().faco_engine_set_local_search(self._handle, _local_search_mode(mode)))

    def set_roulette_margin(self, margin: float) -> None:
        """Test hook: widen the fast roulette's margin (float('inf'): the exact
        left-to-right sums at every step). Results must not change."""
        check(lib().faco_engine_set_roulette_margin(self._handle, float(margin)))

    # Learned edge prior (this engine's extension; gplsaco_amd.prior).
    def _instance_n(self, instance_key) -> int:
        """City count of a registered instance (the engine's dimension)."""
        return self.dimension

    def edge_features(self, instance_key) -> np.ndarray:
        """Features [n * primary_width, 8] of a registered instance's candidate edges."""
        out = np.zeros((self._instance_n(instance_key) * self.settings.primary_width, 8), dtype=np.float32)
        check(lib().faco_engine_edge_features(self._handle, _unsigned(instance_key, "instance_key"), out.reshape(-1)))
        return out

    def set_edge_prior(self, instance_key, multipliers=None) -> None:
        """Heuristic multipliers [n, primary_width] of an instance (None: plain FACO)."""
        key = _unsigned(instance_key, "instance_key")
        if multipliers is None:
            check(lib().faco_engine_set_edge_prior(self._handle, key, None))
            return
        m = np.ascontiguousarray(multipliers, dtype=np.float32).reshape(-1)
        if m.size != self._instance_n(key) * self.settings.primary_width:
            raise ValueError("multipliers must be [n, primary_width]")
        check(lib().faco_engine_set_edge_prior(self._handle, key, m.ctypes.data_as(ctypes.c_void_p)))

    def apply_edge_prior(self, prior, instance_key, exponent=None) -> np.ndarray:
        """Score the instance's candidate edges with `prior` (an EdgePrior) and use
        sigmoid(logit) ** exponent as their heuristic multipliers (exponent
        None: the prior's default, 2); returns them."""
        from .prior import PRIOR_EXPONENT
        exponent = PRIOR_EXPONENT if exponent is None else float(exponent)
        out = np.zeros((self._instance_n(instance_key), self.settings.primary_width), dtype=np.float32)
        check(lib().faco_engine_apply_edge_prior(self._handle, prior._handle, _unsigned(instance_key, "instance_key"),
                                                 out.ctypes.data_as(ctypes.c_void_p), exponent))
        return out

    def autotune(self, instance_keys, seeds, iterations: int = 4) -> dict:
        """Time every ant-kernel launch shape that fits (ants per workgroup,
        neighbour rows in LDS or not) for `iterations` iterations of plain
        FACO on these tasks and keep the fastest. Results of evaluations do
        not depend on the shape; only their speed does."""
        keys, seeds = _keys(instance_keys, "instance_keys"), _keys(seeds, "seeds")
        if keys.size != seeds.size:
            raise ValueError("autotune needs complete task arrays")
        table = (TuneEntry * 32)()
        count = ctypes.c_uint32()
        check(lib().faco_engine_autotune(self._handle, keys, seeds, keys.size, _unsigned(iterations, "iterations", 32),
                                         table, 32, ctypes.byref(count)))
        rows = [{"waves": t.waves, "neighbour_rows_in_lds": bool(t.nbr_in_lds),
                 "blocks_per_colony": t.blocks_per_colony, "ms_per_iteration": t.ms_per_iteration}
                for t in table[:min(count.value, 32)]]
        best = min(rows, key=lambda r: r["ms_per_iteration"])
        return {"table": rows, "best": best, "launch_shape": self.launch_shape(keys.size)}

    def autotune_control(self, instance_keys, seeds, mode, *, programs=(), controllers=(), baseline=None,
                         variant=3, experiment_mask=0xFFFFFFFF, iterations: int = 2) -> dict:
        """The autotuner on a controlled run itself: every launch shape that
        fits is timed on `iterations` batches of this controlled run over the
        panel (colony d * panel + t under decider d, as in
        evaluate_control_raw); the fastest is kept. "persistent" marks the
        shapes that run the persistent controlled kernel."""
        keys, seeds = _keys(instance_keys, "instance_keys"), _keys(seeds, "seeds")
        if keys.size != seeds.size:
            raise ValueError("autotune needs complete task arrays")
        spec, prog_arr, ctrl_arr = _control_spec(mode, programs, controllers, baseline, variant, experiment_mask, False)
        table = (TuneEntry * 32)()
        count = ctypes.c_uint32()
        check(lib().faco_engine_autotune_control(self._handle, keys, seeds, keys.size, ctypes.byref(spec),
                                                 _unsigned(iterations, "iterations", 32), table, 32,
                                                 ctypes.byref(count)))
        del prog_arr, ctrl_arr
        rows = [{"waves": t.waves, "neighbour_rows_in_lds": bool(t.nbr_in_lds),
                 "blocks_per_colony": t.blocks_per_colony, "persistent": bool(t.pad),
                 "ms_per_iteration": t.ms_per_iteration}
                for t in table[:min(count.value, 32)]]
        best = min(rows, key=lambda r: r["ms_per_iteration"])
        return {"table": rows, "best": best}

    def set_waves_per_block(self, waves: int, blocks_per_colony: int = 0) -> None:
        """Force the ant kernel's shape: waves (ants in flight) per workgroup,
        1..16, plus 0x100 to keep neighbour rows out of LDS; workgroups per
        colony (0: automatic). waves = 0 restores the automatic shape."""
        check(lib().faco_engine_set_waves(self._handle, _unsigned(waves, "waves", 32),
                                          _unsigned(blocks_per_colony, "blocks_per_colony", 32)))

    def register_problem(self, instance_key, coordinates) -> dict:
        if type(instance_key) is not int:
            raise ValueError("instance key must be an integer")
        key = _unsigned(instance_key, "instance_key")
        xy = _coordinates(coordinates)
        cheap, prep = ctypes.c_double(), ctypes.c_double()
        check(lib().faco_engine_register_problem(self._handle, key, xy.reshape(-1), xy.shape[0], ctypes.byref(cheap),
                                                 ctypes.byref(prep)))
        return {"cheap_seconds": cheap.value, "preparation_seconds": prep.value}

    def register_graph_problem(self, instance_key, coordinates, dictionary) -> dict:
        """Hard-mode registration (reference bindings.cpp:574-622): the spec
        holds graph_spec_id 1, common_initial_tour, edges [[a, b], ...] and the
        primary, backup and ls rows ([n][width] lists, right-padded with n),
        which replace the kNN rows; the initial tour must be the prepared
        one."""
        xy = _coordinates(coordinates)
        fields = ("graph_spec_id", "common_initial_tour", "edges", "primary", "backup", "ls")
        if not isinstance(dictionary, dict) or set(dictionary) != set(fields):
            raise ValueError("graph registration fields missing or extra")
        if type(dictionary["graph_spec_id"]) is not int or dictionary["graph_spec_id"] != 1:
            raise ValueError("unknown graph spec version")
        n, s = xy.shape[0], self.settings

        def rows(name, width):
            value = np.asarray(dictionary[name], dtype=np.int64)
            if value.shape != (n, width) or (value.size and (value.min() < 0 or value.max() > n)):
                raise ValueError(f"graph rows '{name}' must be [n][{width}] cities (padding n)")
            return np.ascontiguousarray(value, dtype=np.uint32).reshape(-1)
        tour = np.asarray(dictionary["common_initial_tour"], dtype=np.int64)
        if tour.shape != (n,) or tour.min() < 0 or tour.max() >= n:
            raise ValueError("common initial tour must hold n cities")
        edges = np.asarray(dictionary["edges"], dtype=np.int64).reshape(-1, 2) if len(dictionary["edges"]) else \
            np.zeros((0, 2), dtype=np.int64)
        if edges.size and (edges.min() < 0 or edges.max() >= n):
            raise ValueError("graph edge out of range")
        edges = np.ascontiguousarray(edges, dtype=np.uint32)
        primary, ls = rows("primary", s.primary_width), rows("ls", s.ls_width)
        backup = rows("backup", s.backup_width) if s.backup_width else None
        cheap, prep = ctypes.c_double(), ctypes.c_double()
        key = _unsigned(instance_key, "instance_key")
        check(lib().faco_engine_register_graph_problem(
            self._handle, key, xy.reshape(-1), n,
            edges.ctypes.data_as(ctypes.c_void_p), edges.shape[0], np.ascontiguousarray(tour, dtype=np.uint32), primary,
            backup.ctypes.data_as(ctypes.c_void_p) if backup is not None else None, ls, ctypes.byref(cheap),
            ctypes.byref(prep)))
        self._graph_edges[key] = int(edges.shape[0])
        return {"cheap_seconds": cheap.value, "preparation_seconds": prep.value}

    def _escape_flag(self) -> int:
        if self.disable_escape and self.constraint_mode != "escape":
            raise ValueError("only an Escape-mode engine can disable its escape")
        return 16 if self.disable_escape else 0

    def _constraint_output(self, out: dict, keys, rejections: int) -> dict:
        """The Hard- and Escape-mode keys of a result (reference
        bindings.cpp:364-388): the graph's undirected edge count per colony,
        the 2-opt moves rejected for leaving the graph and, for Escape, the
        escape counters summed over the colonies."""
        if self.constraint_mode != "unrestricted":
            reserved = 0
            if self.constraint_mode == "escape":
                counters, reserved = self._escape_counters(len(keys))
                out.update({"escape_spec_id": 1, "escape_edge_capacity_per_ant": 64,
                            "escape_counters": {name: int(v) for name, v in zip(ESCAPE_COUNTERS, counters.sum(axis=0))}})
            out.update({"constraint_mode": self.constraint_mode, "graph_spec_id": 1,
                        "graph_edges_per_colony": [self._graph_edges[int(k)] for k in keys],
                        "completed_constraint_rejections": int(rejections),
                        "preparation_scope": "engine_only; external candidate graph cached",
                        "reserved_escape_device_bytes": int(reserved)})
        return out

    def _escape_counters(self, tasks):
        out = np.zeros((_unsigned(tasks, "tasks", 32), len(ESCAPE_COUNTERS)), dtype=np.uint64)
        reserved = ctypes.c_uint64()
        check(lib().faco_engine_escape_counters(self._handle, out.reshape(-1), out.shape[0], ctypes.byref(reserved)))
        return out, reserved.value

    def colony_escape_counters(self, tasks) -> np.ndarray:
        """Escape mode: each task's escape counters of the last count
        evaluation, [tasks][11] in ESCAPE_COUNTERS order (the reference's
        EscapeStats, escape_slots.hpp:70-75)."""
        return self._escape_counters(tasks)[0]

    def colony_constraint_rejections(self, tasks) -> np.ndarray:
        """Hard mode: the 2-opt moves each task's colony rejected in the last
        count evaluation (its own batches for a continuation), in task order."""
        out = np.zeros(_unsigned(tasks, "tasks", 32), dtype=np.uint64)
        check(lib().faco_engine_constraint_rejections(self._handle, out, out.size))
        return out

    def register_problems_device(self, instance_keys, coordinates) -> dict:
        """Register many instances with the preparation on the GPU (the same
        rows, initial tour and cost as register_problem). coordinates: one
        FP64 [n, 2] array per key. Returns the wall time of the call."""
        keys = [_unsigned(k, "instance_key") for k in instance_keys]
        arrays = [_coordinates(xy) for xy in coordinates]
        if len(keys) != len(arrays) or not keys:
            raise ValueError("one coordinate array per instance key")
        sizes = np.array([a.shape[0] for a in arrays], dtype=np.uint32)
        xy = np.ascontiguousarray(np.concatenate(arrays), dtype=np.float64)
        seconds = ctypes.c_double()
        check(lib().faco_engine_register_problems_device(self._handle, np.array(keys, dtype=np.uint64), xy.reshape(-1),
                                                         ptr(sizes), len(keys), ctypes.byref(seconds)))
        return {"seconds": seconds.value, "instances": len(keys)}

    def preparation_profile(self, instance_key) -> dict:
        out = np.zeros(4, dtype=np.float64)
        check(lib().faco_engine_preparation_profile(self._handle, _unsigned(instance_key, "instance_key"), out))
        return {"candidates_and_scales_seconds": out[0], "nearest_neighbor_seconds": out[1],
                "initial_ls_seconds": out[2], "finalization_seconds": out[3]}

    def set_preparation_charges(self, instance_key, cheap_seconds, preparation_seconds) -> None:
        key = _unsigned(instance_key, "instance_key")
        for v in (cheap_seconds, preparation_seconds):
            if type(v) not in (int, float) or not np.isfinite(v) or v < 0:
                raise ValueError("frozen charges must be finite non-negative numbers")
        check(lib().faco_engine_set_preparation_charges(self._handle, key, float(cheap_seconds),
                                                        float(preparation_seconds)))
        self._charges[key] = (float(cheap_seconds), float(preparation_seconds))

    def problem_data(self, instance_key) -> dict:
        n, s = self._instance_n(instance_key), self.settings
        primary = np.zeros(n * s.primary_width, dtype=np.uint32)
        backup = np.zeros(max(1, n * s.backup_width), dtype=np.uint32)
        ls = np.zeros(n * s.ls_width, dtype=np.uint32)
        tour = np.zeros(n, dtype=np.uint32)
        cost = ctypes.c_int64()
        check(lib().faco_engine_problem_data(self._handle, _unsigned(instance_key, "instance_key"), ptr(primary),
                                             ptr(backup), ptr(ls), ptr(tour), ctypes.byref(cost)))
        return {"primary": primary.reshape(n, -1), "backup": backup[: n * s.backup_width].reshape(n, -1),
                "ls": ls.reshape(n, -1), "initial_tour": tour, "initial_cost": int(cost.value)}

    def evaluate_faco_evaluations(self, instance_keys, seeds, evaluation_limit_per_colony, mne=8,
                                  preparation_mode="cached", *, checkpoint_iterations=(), profile_events_only=False,
                                  use_graph=True, profile_kernels=False, time_kernels=False,
                                  continue_run=False) -> dict:
        """Plain FACO for evaluation_limit_per_colony tours per colony
        (reference bindings.cpp:701-720). continue_run (an engine extension):
        resume the previous plain evaluation of the same tasks and MNE target
        from its last batch instead of resetting the colonies, so that a run
        and its continuation equal one longer run; work and evaluations then
        count the continuation only."""
        keys, seeds = _keys(instance_keys, "instance_keys"), _keys(seeds, "seeds")
        if keys.size != seeds.size:
            raise ValueError("FACO count entry needs complete task arrays")
        if preparation_mode not in ("cached", "end_to_end"):
            raise ValueError("unknown preparation mode")
        limit = _unsigned(evaluation_limit_per_colony, "evaluation_limit_per_colony")
        mne = _unsigned(mne, "mne", 32)
        tasks = keys.size
        request = _CheckpointRequest(checkpoint_iterations, tasks, self.dimension)
        tours = np.zeros((tasks, self.dimension), dtype=np.uint32)
        costs = np.zeros(tasks, dtype=np.int64)
        work = np.zeros((tasks, 2), dtype=np.uint64)
        result = Result()
        started = time.perf_counter()
        flags = (1 if use_graph else 0) | (2 if profile_kernels else 0) | (4 if time_kernels or profile_events_only else 0) \
            | (8 if continue_run else 0) | self._escape_flag()
        check(lib().faco_engine_evaluate_faco(self._handle, keys, seeds, tasks, limit, mne, flags,
                                              ctypes.byref(result), tours.reshape(-1), costs, work.reshape(-1),
                                              request.argument()))
        out = _batch_output(result, tours, costs, preparation_mode, time.perf_counter() - started)
        out["colony_work"] = work
        self._constraint_output(out, keys, result.completed_constraint_rejections)
        if request.points:
            out["checkpoints"] = request.output()
        if profile_events_only:
            out["production_kernel_milliseconds"] = _production_stages(result.kernel_milliseconds, False)
        if profile_kernels or time_kernels:
            out["kernel_milliseconds"] = {"construction_and_search": result.kernel_milliseconds[0],
                                          "reduction_and_pheromone": result.kernel_milliseconds[1]}
        if profile_kernels:
            out["ant_phase_milliseconds"] = {"setup": result.kernel_milliseconds[2],
                                             "construction": result.kernel_milliseconds[3],
                                             "local_search": result.kernel_milliseconds[4]}
            km = result.kernel_milliseconds
            out["ant_subphase_megaclocks"] = {"select": km[5], "relocate": km[6], "mne_push": km[7],
                                              "ls_scan": km[8], "ls_apply": km[9]}
            # per-ant milliseconds from taking the ant to its write-back, and the
            # share of the resident waves' time spent holding an ant (both summed
            # over iterations)
            out["ant_hold_milliseconds"] = km[10]
            out["wave_utilisation"] = km[11]
        return out


    # ------------------------------------------------------------------
    # Time-budget evaluations (reference bindings.cpp:662-700, 937-971;
    # cpp/include/gp_faco/deadline.hpp)

    def _deadline(self, instance_keys, seeds, budget_seconds, mode, *, mne=0, program=None,
                  experiment_mask=0xFFFFFFFF, fixed_batches=0, fixed_ratio=-1.0, profile=False) -> dict:
        keys, seeds = _keys(instance_keys, "instance_keys"), _keys(seeds, "seeds")
        if keys.size != seeds.size or type(budget_seconds) not in (int, float):
            raise ValueError("task arrays or budget type invalid")
        if mode not in ("cached_charged", "end_to_end"):
            raise ValueError("unknown preparation mode")
        spec = DeadlineSpec()
        spec.budget_seconds = float(budget_seconds)
        spec.mne_target = mne
        spec.preparation_mode = 0 if mode == "cached_charged" else 1
        spec.experiment_mask = experiment_mask
        spec.fixed_batches = fixed_batches
        spec.fixed_elapsed_ratio = fixed_ratio
        spec.profile = 1 if profile else 0
        if program is not None:
            spec.program = ctypes.addressof(program)
        tasks, n = keys.size, self.dimension
        tours = np.zeros((tasks, n), dtype=np.uint32)
        costs = np.zeros(tasks, dtype=np.float64)
        has = np.zeros(tasks, dtype=np.uint8)
        completed = np.zeros(tasks, dtype=np.float64)
        states = (ControlState * tasks)()
        capacity = fixed_batches or 1 << 16
        ratios = np.zeros(capacity, dtype=np.float64)
        rows = (BatchProfile * max(fixed_batches, 1))()
        r = DeadlineResult()
        check(lib().faco_engine_evaluate_deadline(
            self._handle, keys, seeds, tasks, ctypes.byref(spec), ctypes.byref(r), tours.reshape(-1), costs, has,
            completed, ctypes.addressof(states) if program is not None else None,
            ratios.ctypes.data_as(ctypes.c_void_p), capacity, ctypes.addressof(rows) if profile else None))
        items = [{"has_incumbent": bool(has[c]), "tour": tours[c].tolist() if has[c] else [],
                  "cost": float(costs[c]) if has[c] else None, "completed_seconds": float(completed[c])}
                 for c in range(tasks)]
        out = {"items": items, "budget_seconds": r.budget_seconds, "elapsed_seconds": r.elapsed_seconds,
               "actual_seconds": r.actual_seconds, "charged_seconds": r.charged_seconds,
               "last_batch_completed_seconds": r.last_batch_completed_seconds, "overrun_seconds": r.overrun_seconds,
               "launched_batches": r.launched_batches, "completed_batches": r.completed_batches,
               "discarded_batches": r.discarded_batches,
               "completed_construction_steps": r.completed_construction_steps,
               "completed_ls_evaluations": r.completed_ls_evaluations,
               "allocated_device_bytes": r.allocated_device_bytes,
               "preparation_completed": bool(r.preparation_completed), "preparation_mode": mode,
               "progress_features": ratios[:min(r.launched_batches, capacity)].tolist()}
        if program is not None:
            out["control_states"] = [_control_state_dict(x) for x in states]
        if profile:
            stages = ("initialization", "features_and_regions", "gp_scoring", "action_and_restart",
                      "construction_and_ls", "reduction", "pheromone", "ant_fingerprints", "archive_and_feedback")
            out["profile"] = {
                "profile_version": 1, "setup_seconds": r.setup_seconds, "host_pack_seconds": 0.0,
                "upload_call_seconds": r.upload_seconds,
                "initialization_gpu_milliseconds": r.initialization_gpu_milliseconds,
                "initialization_includes_uploads": False, "initialization_download_seconds": 0.0,
                "diagnostic_device_bytes": 0,
                "batches": [{"batch": row.batch, "committed": bool(row.committed), "wall_seconds": row.wall_seconds,
                             "download_seconds": row.download_seconds,
                             "verification_seconds": row.verification_seconds,
                             "collection_seconds": row.collection_seconds,
                             "gpu_milliseconds": dict(zip(stages, row.gpu_milliseconds)),
                             "ant_phase_cycles": []}
                            for row in rows[:r.profiled_batches]]}
        return out

    def evaluate(self, instance_keys, seeds, budget_seconds, mne_target, preparation_mode="cached_charged") -> dict:
        """Plain FACO under a time budget (reference bindings.cpp:662-680):
        every colony's incumbent is the best tour offered within the budget
        (cheap tour, prepared tour, then the global best after each batch
        that completes in time)."""
        if type(mne_target) is not int:
            raise ValueError("task arrays or budget/MNE type invalid")
        mne = _unsigned(mne_target, "mne_target", 32)
        return self._deadline(instance_keys, seeds, budget_seconds, preparation_mode, mne=mne)

    def evaluate_program(self, instance_keys, seeds, budget_seconds, program, preparation_mode="cached_charged",
                         experiment_mask=0xFFFFFFFF) -> dict:
        """A feature-spec-1 program controls the colonies under a time budget
        (reference bindings.cpp:681-700): its progress feature is the
        elapsed share of the budget at each decision."""
        if type(experiment_mask) is not int:
            raise ValueError("task arrays or budget/mask type invalid")
        p = read_program(program)
        mask = _unsigned(experiment_mask, "experiment_mask", 32)
        return self._deadline(instance_keys, seeds, budget_seconds, preparation_mode, program=p, experiment_mask=mask)

    def run_program_diagnostic(self, instance_keys, seeds, program, batches, elapsed_ratio=0.5, profile=True,
                               experiment_mask=0xFFFFFFFF, preparation_mode="cached_charged",
                               budget_seconds=120.0) -> dict:
        """Fixed-batch program run for diagnostics, not fitness (reference
        bindings.cpp:937-971): `batches` batches (1..10000) with a constant
        progress feature, per-batch HIP-event and host timings when
        `profile`, inside a time budget."""
        if type(batches) is not int or type(profile) is not bool or type(elapsed_ratio) not in (int, float) \
                or type(experiment_mask) is not int or type(budget_seconds) not in (int, float):
            raise ValueError("diagnostic arrays or timing argument types invalid")
        if not 1 <= batches <= 10000:
            raise ValueError("diagnostic batch count must be in 1..10000")
        p = read_program(program)
        mask = _unsigned(experiment_mask, "experiment_mask", 32)
        out = self._deadline(instance_keys, seeds, budget_seconds, preparation_mode, program=p, experiment_mask=mask,
                             fixed_batches=batches, fixed_ratio=float(elapsed_ratio), profile=profile)
        out["scope"] = "fixed_batch_diagnostic_not_fitness"
        out.setdefault("profile", {"profile_version": 1, "batches": []})
        return out

    # ------------------------------------------------------------------
    # Controlled evaluations (reference bindings.cpp:721-879)

    def evaluate_control_raw(self, instance_keys, seeds, evaluation_limit_per_colony, mode, *, programs=(),
                             controllers=(), baseline=None, variant=3, experiment_mask=0xFFFFFFFF,
                             trace_batches=(), use_graph=True, force_fingerprint_collisions=False,
                             checkpoint_iterations=(), profile_events_only=False) -> dict:
        """Controlled run through faco_engine_evaluate_control; arrays out.

        mode: "program", "controller", "baseline" or "factorial"; colony
        c = d * panel + t runs task t under decider d. trace_batches
        (ascending, 0-based) selects the batches whose decisions are kept;
        checkpoint_iterations records the global-best tours after those
        numbers of completed batches; profile_events_only times each launch
        with HIP events (eager launches, same results)."""
        keys, seeds = _keys(instance_keys, "instance_keys"), _keys(seeds, "seeds")
        if keys.size != seeds.size:
            raise ValueError("controlled entry points need complete task arrays")
        limit = _unsigned(evaluation_limit_per_colony, "evaluation_limit_per_colony")

        def call(spec, flags, trace, result, tours, costs, work, states, checkpoints):
            return lib().faco_engine_evaluate_control(self._handle, keys, seeds, keys.size, limit, spec, flags, trace,
                                                      result, tours, costs, work, states, checkpoints)
        return self._control_run(keys, mode, call, programs=programs, controllers=controllers,
                                 baseline=baseline, variant=variant, experiment_mask=experiment_mask,
                                 trace_batches=trace_batches, use_graph=use_graph,
                                 force_fingerprint_collisions=force_fingerprint_collisions,
                                 checkpoint_iterations=checkpoint_iterations, profile_events_only=profile_events_only)

    def _control_run(self, panel_keys, mode, call, *, programs=(), controllers=(), baseline=None, variant=3,
                     experiment_mask=0xFFFFFFFF, trace_batches=(), use_graph=True, force_fingerprint_collisions=False,
                     checkpoint_iterations=(), profile_events_only=False) -> dict:
        """One controlled native call over the tasks of `panel_keys` (their
        instance keys): `call` receives the spec, flags, trace and output
        pointers of faco_engine_evaluate_control."""
        panel = len(panel_keys)
        spec, prog_arr, ctrl_arr = _control_spec(mode, programs, controllers, baseline, variant, experiment_mask,
                                                 force_fingerprint_collisions)
        C = panel * spec.deciders
        tours = np.zeros((C, self.dimension), dtype=np.uint32)
        costs = np.zeros(C, dtype=np.int64)
        work = np.zeros((C, 2), dtype=np.uint64)
        states = (ControlState * C)()
        result = Result()
        request = _CheckpointRequest(checkpoint_iterations, C, self.dimension)
        batches = np.ascontiguousarray(sorted(trace_batches), dtype=np.uint32)
        k = batches.size
        traces = {}
        trace = None
        if k:
            traces = {"features": np.zeros((k, 12, C, 32), dtype=np.float32),
                      "scores": np.zeros((k, C, 32), dtype=np.float32),
                      "masks": np.zeros((k, C), dtype=np.uint32), "actions": np.zeros((k, C), dtype=np.int32),
                      "alternatives": np.zeros((k, C), dtype=np.uint32),
                      "regions": np.zeros((k, C, 2, 4, 16), dtype=np.uint32),
                      "costs": np.zeros((k, C, 3), dtype=np.int64)}
            trace = ControlTrace()
            trace.batches = batches.ctypes.data_as(ctypes.c_void_p)
            trace.count = k
            for name, array in traces.items():
                setattr(trace, name, array.ctypes.data_as(ctypes.c_void_p))
        started = time.perf_counter()
        check(call(ctypes.byref(spec), (1 if use_graph else 0) | (4 if profile_events_only else 0) | self._escape_flag(),
                   ctypes.byref(trace) if trace else None, ctypes.byref(result), tours.reshape(-1), costs,
                   work.reshape(-1), ctypes.addressof(states), request.argument()))
        del prog_arr, ctrl_arr
        rejections = None
        if self.constraint_mode != "unrestricted":
            rejections = self.colony_constraint_rejections(C)
        return {"result": result, "tours": tours, "costs": costs, "work": work, "elapsed": time.perf_counter() - started,
                "keys": np.resize(panel_keys, C), "rejections": rejections,
                "checkpoints": request.output() if request.points else None,
                "profile_events_only": bool(profile_events_only),
                "states": [_control_state_dict(st) for st in states],
                **{"trace_" + name: array for name, array in traces.items()}}

    @staticmethod
    def _decision_batches(decision_iterations, limit, ants):
        points = list(decision_iterations)
        if points and (any(type(v) is not int for v in points) or points[0] <= 0 or points[-1] > limit // ants
                       or sorted(set(points)) != points):
            raise ValueError("decision iterations must be increasing iteration numbers inside the budget")
        return [v - 1 for v in points]

    def _control_output(self, raw, decision_iterations, kind) -> dict:
        out = _batch_output(raw["result"], raw["tours"], raw["costs"], "cached", raw["elapsed"])
        out["control_states"] = raw["states"]
        self._constraint_output(out, raw["keys"], raw["result"].completed_constraint_rejections)
        if raw.get("checkpoints"):
            out["checkpoints"] = raw["checkpoints"]
        if raw.get("profile_events_only"):
            out["production_kernel_milliseconds"] = _production_stages(raw["result"].kernel_milliseconds, True)
        if decision_iterations:
            rows = []
            C = raw["tours"].shape[0]
            for k, iteration in enumerate(decision_iterations):
                for colony in range(C):
                    row = {"iteration": iteration, "colony": colony,
                           "legal_mask": int(raw["trace_masks"][k, colony]),
                           "action": int(raw["trace_actions"][k, colony]),
                           "features": raw["trace_features"][k, :, colony, :].tolist()}
                    scores = raw["trace_scores"][k, colony]
                    if kind == "conditional":
                        row["stages"] = _controller_stages(scores, row["legal_mask"], row["action"])
                    elif kind != "fixed":
                        row["scores"] = [float(v) if np.isfinite(v) else None for v in scores]
                    rows.append(row)
            out["decisions"] = rows
        return out

    def evaluate_program_evaluations(self, instance_keys, seeds, evaluation_limit_per_colony, program,
                                     preparation_mode="cached", experiment_mask=0xFFFFFFFF, *, record_behavior=False,
                                     checkpoint_iterations=(), profile=False, profile_events_only=False,
                                     decision_iterations=(), use_graph=True) -> dict:
        if type(evaluation_limit_per_colony) is not int or type(experiment_mask) is not int:
            raise ValueError("count entry needs an integer limit and mask")
        if preparation_mode not in ("cached", "end_to_end"):
            raise ValueError("count entry takes cached or end_to_end preparation")
        if record_behavior or profile:
            raise ValueError("behaviour rows and profiles are not available in this engine")
        p = read_program(program)
        batches = self._decision_batches(decision_iterations, evaluation_limit_per_colony, self.settings.ants)
        raw = self.evaluate_control_raw(instance_keys, seeds, evaluation_limit_per_colony, "program", programs=[p],
                                        experiment_mask=experiment_mask, trace_batches=batches, use_graph=use_graph,
                                        checkpoint_iterations=checkpoint_iterations,
                                        profile_events_only=profile_events_only)
        return self._control_output(raw, list(decision_iterations), "program")

    def evaluate_controller_evaluations(self, instance_keys, seeds, evaluation_limit_per_colony, controller,
                                        experiment_mask=0xFFFFFFFF, *, decision_iterations=(), use_graph=True) -> dict:
        c = read_controller(controller)
        batches = self._decision_batches(decision_iterations, evaluation_limit_per_colony, self.settings.ants)
        raw = self.evaluate_control_raw(instance_keys, seeds, evaluation_limit_per_colony, "controller",
                                        controllers=[c], experiment_mask=experiment_mask, trace_batches=batches,
                                        use_graph=use_graph)
        return self._control_output(raw, list(decision_iterations), "conditional" if c.kind else "program")

    def evaluate_baseline_evaluations(self, instance_keys, seeds, evaluation_limit_per_colony, policy,
                                      preparation_mode="cached", experiment_mask=0xFFFFFFFF, *, record_behavior=False,
                                      checkpoint_iterations=(), profile=False, profile_events_only=False,
                                      decision_iterations=(), use_graph=True) -> dict:
        if preparation_mode not in ("cached", "end_to_end"):
            raise ValueError("count entry takes cached or end_to_end preparation")
        if record_behavior or profile:
            raise ValueError("behaviour rows and profiles are not available in this engine")
        b = read_baseline(policy)
        batches = self._decision_batches(decision_iterations, evaluation_limit_per_colony, self.settings.ants)
        raw = self.evaluate_control_raw(instance_keys, seeds, evaluation_limit_per_colony, "baseline", baseline=b,
                                        experiment_mask=_unsigned(experiment_mask, "experiment_mask", 32),
                                        trace_batches=batches, use_graph=use_graph,
                                        checkpoint_iterations=checkpoint_iterations,
                                        profile_events_only=profile_events_only)
        out = self._control_output(raw, list(decision_iterations), "fixed")
        out["baseline_policy"] = baseline_output(b)
        return out

    # ------------------------------------------------------------------
    # State snapshots and paired forks (reference bindings.cpp:880-934,
    # batch_engine.cu:465-490)

    @staticmethod
    def _state_decider(kind, argument) -> dict:
        if kind == "program":
            return {"programs": [read_program(argument)]}
        return {"baseline": read_baseline(argument)}

    def _capture_state(self, kind, instance_keys, seeds, source_evaluations, capture_after, argument,
                       experiment_mask, decision_iterations=()) -> dict:
        keys, seeds = _keys(instance_keys, "instance_keys"), _keys(seeds, "seeds")
        if keys.size != seeds.size:
            raise ValueError("a capture needs complete task arrays")
        source = _unsigned(source_evaluations, "source_evaluation_limit_per_colony")
        after = _unsigned(capture_after, "capture_after_evaluations")
        mask = _unsigned(experiment_mask, "experiment_mask", 32)
        decider = self._state_decider(kind, argument)
        handle = ctypes.c_void_p()

        def call(spec, flags, trace, result, tours, costs, work, states, checkpoints):
            return lib().faco_engine_capture_control(self._handle, keys, seeds, keys.size, source, after, spec, flags,
                                                     trace, result, tours, costs, work, states, ctypes.byref(handle))
        batches = self._decision_batches(decision_iterations, after, self.settings.ants)
        raw = self._control_run(keys, kind, call, experiment_mask=mask, trace_batches=batches, **decider)
        state = CountedState(handle)
        evaluation = self._control_output(raw, list(decision_iterations), "program" if kind == "program" else "fixed")
        if kind == "baseline":
            evaluation["baseline_policy"] = baseline_output(decider["baseline"])
        return {"evaluation": evaluation, "state": state}

    def _continue_state(self, kind, state, additional_evaluations, argument, fork=None, decision_iterations=()) -> dict:
        if not isinstance(state, CountedState):
            raise TypeError("state must be a CountedState")
        additional = _unsigned(additional_evaluations, "additional_evaluations_per_colony")
        decider = self._state_decider(kind, argument)

        def call(spec, flags, trace, result, tours, costs, work, states, checkpoints):
            return lib().faco_engine_continue_control(self._handle, state._handle, additional, spec,
                                                      ctypes.byref(fork) if fork is not None else None, flags, trace,
                                                      result, tours, costs, work, states)
        batches = self._decision_batches(decision_iterations, additional, self.settings.ants)
        tasks = [t["instance_key"] for t in state.describe()["tasks"]]
        raw = self._control_run(np.resize(np.array(tasks, dtype=np.uint64), state.colonies), kind, call,
                                trace_batches=batches, **decider)
        out = self._control_output(raw, list(decision_iterations), "program" if kind == "program" else "fixed")
        if kind == "baseline":
            out["baseline_policy"] = baseline_output(decider["baseline"])
        return out

    def capture_program_state(self, instance_keys, seeds, source_evaluation_limit_per_colony,
                              capture_after_evaluations, program, experiment_mask=0xFFFFFFFF, *,
                              decision_iterations=()) -> dict:
        """Run a program-controlled panel for capture_after_evaluations per
        colony (the progress feature still measured against the source
        budget) and return {"evaluation": ..., "state": CountedState}.
        decision_iterations (1-based batches of this call) keeps those
        decisions as in evaluate_program_evaluations."""
        return self._capture_state("program", instance_keys, seeds, source_evaluation_limit_per_colony,
                                   capture_after_evaluations, program, experiment_mask, decision_iterations)

    def capture_baseline_state(self, instance_keys, seeds, source_evaluation_limit_per_colony,
                               capture_after_evaluations, policy, experiment_mask=0xFFFFFFFF, *,
                               decision_iterations=()) -> dict:
        return self._capture_state("baseline", instance_keys, seeds, source_evaluation_limit_per_colony,
                                   capture_after_evaluations, policy, experiment_mask, decision_iterations)

    def continue_program_state(self, state, additional_evaluations_per_colony, program, *,
                               decision_iterations=()) -> dict:
        """Resume a snapshot for additional evaluations per colony; capture +
        continuation equals the uninterrupted run. decision_iterations count
        the batches of this call from 1."""
        return self._continue_state("program", state, additional_evaluations_per_colony, program,
                                    decision_iterations=decision_iterations)

    def continue_baseline_state(self, state, additional_evaluations_per_colony, policy, fork_seed=None, region=0,
                                mne=2, *, decision_iterations=()) -> dict:
        """Resume under a baseline policy; with fork_seed the first decision
        is forced to region * 4 + (0 for MNE 2, 3 for MNE 16) and the keys
        are re-derived from the seed (a paired intervention)."""
        fork = Fork()
        fork.enabled = 0 if fork_seed is None else 1
        fork.seed = 0 if fork_seed is None else _unsigned(fork_seed, "fork_seed")
        fork.region = _unsigned(region, "region", 32)
        fork.mne = _unsigned(mne, "mne", 32)
        return self._continue_state("baseline", state, additional_evaluations_per_colony, policy, fork,
                                    decision_iterations)

    def evaluate_factorial_evaluations(self, instance_keys, seeds, evaluation_limit_per_colony, program, policy,
                                       preparation_mode="cached", experiment_mask=0xFFFFFFFF, *,
                                       record_behavior=False, use_graph=True) -> dict:
        if preparation_mode not in ("cached", "end_to_end"):
            raise ValueError("count entry takes cached or end_to_end preparation")
        if record_behavior:
            raise ValueError("behaviour rows are not available in this engine")
        p = read_program(program)
        variant, b = read_factorial(policy)
        raw = self.evaluate_control_raw(instance_keys, seeds, evaluation_limit_per_colony, "factorial", programs=[p],
                                        baseline=b, variant=variant,
                                        experiment_mask=_unsigned(experiment_mask, "experiment_mask", 32),
                                        use_graph=use_graph)
        out = self._control_output(raw, [], "program")
        out["factorial_policy"] = factorial_output(variant, b)
        return out

    def evaluate_population_evaluations(self, instance_keys, seeds, evaluation_limit_per_colony, programs,
                                        experiment_mask=0xFFFFFFFF, *, factorial_policy=None,
                                        profile_events_only=False, use_graph=True) -> list:
        """One launch for the whole population (reference batch_engine.cu:376-460):
        the panel is replicated per individual; per-individual results share
        the batch time equally."""
        keys, seeds = _keys(instance_keys, "instance_keys"), _keys(seeds, "seeds")
        if keys.size != seeds.size:
            raise ValueError("population entry needs a complete shared panel")
        if not isinstance(programs, (list, tuple)) or not programs:
            raise ValueError("population needs at least one program")
        mask = _unsigned(experiment_mask, "experiment_mask", 32)
        controllers = [read_controller(d) for d in programs]
        conditional = any(c.kind for c in controllers)
        variant, b = (read_factorial(factorial_policy) if factorial_policy is not None else (3, None))
        if factorial_policy is not None and len(programs) > 1 and variant not in (1, 2):
            raise ValueError("population factorial runs take M10 or M01 only")
        if conditional:
            if factorial_policy is not None:
                raise ValueError("controller populations take no factorial policy")
            raw = self.evaluate_control_raw(keys, seeds, evaluation_limit_per_colony, "controller",
                                            controllers=controllers, experiment_mask=mask, use_graph=use_graph,
                                            profile_events_only=profile_events_only)
        elif factorial_policy is not None:
            raw = self.evaluate_control_raw(keys, seeds, evaluation_limit_per_colony, "factorial",
                                            programs=[c.trees[0] for c in controllers], baseline=b, variant=variant,
                                            experiment_mask=mask, use_graph=use_graph,
                                            profile_events_only=profile_events_only)
        else:
            raw = self.evaluate_control_raw(keys, seeds, evaluation_limit_per_colony, "program",
                                            programs=[c.trees[0] for c in controllers], experiment_mask=mask,
                                            use_graph=use_graph, profile_events_only=profile_events_only)
        count, panel = len(programs), keys.size
        r = raw["result"]
        outputs = []
        for individual in range(count):
            sl = slice(individual * panel, (individual + 1) * panel)
            work = raw["work"][sl]
            item = _batch_output(r, raw["tours"][sl], raw["costs"][sl], "cached", raw["elapsed"] / count)
            item.update({
                "population_size": count, "population_actual_seconds": r.actual_seconds,
                "time_accounting": "equal_share_of_population_batch",
                "actual_seconds": r.actual_seconds / count, "elapsed_seconds": r.actual_seconds / count,
                "total_tour_evaluations": r.completed_tour_evaluations_per_colony * panel,
                "completed_construction_steps": int(work[:, 0].sum()),
                "completed_ls_evaluations": int(work[:, 1].sum()),
                "control_states": raw["states"][sl],
            })
            self._constraint_output(item, raw["keys"][sl],
                                    int(raw["rejections"][sl].sum()) if raw["rejections"] is not None else 0)
            for entry in item["items"]:
                entry["completed_seconds"] = raw["elapsed"] / count
            if factorial_policy is not None:
                item["factorial_policy"] = factorial_output(variant, b)
            if profile_events_only:   # one launch sequence for the population: each individual its equal share
                item["production_kernel_milliseconds"] = {
                    k: v / count for k, v in _production_stages(r.kernel_milliseconds, True).items()}
            outputs.append(item)
        return outputs


class MixedFacoBatchEngine(FacoBatchEngine):
    """Instances of different sizes (3 <= n <= max_dimension, equal candidate
    widths) in one engine: every colony runs with its own n and ants (the
    settings' per-size default when ants is 0) in the same kernel launches,
    largest work first; plain FACO and the controlled (GP / baseline)
    evaluations alike. No counterpart in the reference, which needs one n
    per engine (cuda/batch_engine.cu:78-99). The evaluation limit must be a
    whole number of every colony's ant batches (each colony stops at its own
    budget); tours come back at each colony's own n."""

    def __init__(self, max_dimension, colonies, settings, population_size=1):
        if type(max_dimension) is not int or type(colonies) is not int:
            raise ValueError("max_dimension/colonies must be integers")
        if not isinstance(settings, FixedFacoSettings):
            raise TypeError("settings must be FixedFacoSettings")
        population = _unsigned(population_size, "population_size", 32)
        handle = ctypes.c_void_p()
        check(lib().faco_engine_create_mixed(max_dimension, colonies, ctypes.byref(settings._native()), population,
                                             ctypes.byref(handle)))
        self._handle = handle
        self.dimension, self.colonies, self.population_size = max_dimension, colonies, population
        s = Settings()
        check(lib().faco_engine_settings(self._handle, ctypes.byref(s)))
        self.settings = FixedFacoSettings._from_native(s)
        self.constraint_mode = "unrestricted"
        self.disable_escape = False
        self._charges = {}
        self._graph_edges = {}
        self._sizes = {}

    def _instance_n(self, instance_key) -> int:
        key = _unsigned(instance_key, "instance_key")
        if key not in self._sizes:
            raise ValueError("task instance has not been registered")
        return self._sizes[key]

    def register_problem(self, instance_key, coordinates) -> dict:
        out = super().register_problem(instance_key, coordinates)
        self._sizes[_unsigned(instance_key, "instance_key")] = _coordinates(coordinates).shape[0]
        return out

    def register_problems_device(self, instance_keys, coordinates) -> dict:
        out = super().register_problems_device(instance_keys, coordinates)
        for key, xy in zip(instance_keys, coordinates):
            self._sizes[_unsigned(key, "instance_key")] = xy.shape[0]
        return out

    def evaluate_faco_evaluations(self, instance_keys, seeds, evaluation_limit_per_colony, mne=8,
                                  preparation_mode="cached", **kwargs) -> dict:
        out = super().evaluate_faco_evaluations(instance_keys, seeds, evaluation_limit_per_colony, mne,
                                                preparation_mode, **kwargs)
        for key, item in zip(_keys(instance_keys, "instance_keys"), out["items"]):
            item["tour"] = item["tour"][: self._instance_n(int(key))]
        return out

    def _control_output(self, raw, decision_iterations, kind) -> dict:
        out = super()._control_output(raw, decision_iterations, kind)
        for item in out["items"]:   # each row ends at its colony's own n
            tour = item["tour"]
            if 0xFFFFFFFF in tour:
                item["tour"] = tour[: tour.index(0xFFFFFFFF)]
        return out


def _controller_stages(scores, mask, action) -> list:
    """Per-stage view of a conditional controller decision (reference bindings.cpp:77-94)."""
    stages, prefix = [], 0
    for role in range(3):
        width, stride, offset = (2, 16, 0) if role == 0 else (4, 4, 2) if role == 1 else (4, 1, 6)
        legal = [bool((mask >> (prefix + i * stride)) & ((1 << stride) - 1)) for i in range(width)]
        selected = (action - prefix) // stride
        stages.append({"scores": [float(v) for v in scores[offset:offset + width]], "legal": legal,
                       "selected": selected})
        prefix += selected * stride
    return stages


class FixedFacoGpu:
    """Single-instance fixed-iteration FACO (reference fixed_faco_gpu.hpp:44-60)."""

    def __init__(self, coordinates, instance_key, settings):
        xy = _coordinates(coordinates)
        if not isinstance(settings, FixedFacoSettings):
            raise TypeError("settings must be FixedFacoSettings")
        handle = ctypes.c_void_p()
        check(lib().faco_fixed_create(xy.reshape(-1), xy.shape[0], _unsigned(instance_key, "instance_key"),
                                      ctypes.byref(settings._native()), ctypes.byref(handle)))
        self._handle = handle
        self.dimension = xy.shape[0]
        s = Settings()
        check(lib().faco_fixed_settings(self._handle, ctypes.byref(s)))
        self.settings = FixedFacoSettings._from_native(s)

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            lib().faco_fixed_destroy(handle)
            self._handle = None

    def set_local_search(self, mode: str) -> None:
        """'2opt' (the reference's checklist 2-opt, default) or '2opt+oropt'."""
        check(lib().faco_fixed_set_local_search(self._handle, _local_search_mode(mode)))

    def run_iterations(self, seed, batches, mne_target, *, trace=False) -> dict:
        seed = _unsigned(seed, "seed")
        batches = _unsigned(batches, "batches", 32)
        mne = _unsigned(mne_target, "mne_target", 32)
        n, ants, pw = self.dimension, self.settings.ants, self.settings.primary_width
        tour = np.zeros(n, dtype=np.uint32)
        cost, initial = ctypes.c_int64(), ctypes.c_int64()
        steps, evals, seconds = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        tr = {}
        if trace:
            tr = {
                "parent": np.zeros((batches, n), dtype=np.uint32),
                "trails": np.zeros((batches, n * pw), dtype=np.float64),
                "costs": np.zeros((batches, 3), dtype=np.int64),
                "iteration_best": np.zeros(batches, dtype=np.uint32),
                "ants": np.zeros((batches, ants), dtype=_native.ANT_INFO_DTYPE),
                "selected": np.zeros((batches, ants, n), dtype=np.uint32),
                "uniforms": np.zeros((batches, ants, n), dtype=np.float64),
                "tours": np.zeros((batches, ants, n), dtype=np.uint32),
            }
        check(lib().faco_fixed_run(self._handle, seed, batches, mne, tour, ctypes.byref(cost), ctypes.byref(initial),
                                   ctypes.byref(steps), ctypes.byref(evals), ctypes.byref(seconds),
                                   ptr(tr.get("parent")), ptr(tr.get("trails")), ptr(tr.get("costs")),
                                   ptr(tr.get("iteration_best")), ptr(tr.get("ants")), ptr(tr.get("selected")),
                                   ptr(tr.get("uniforms")), ptr(tr.get("tours"))))
        out = {"tour": tour.tolist(), "cost": float(cost.value), "initial_cost": float(initial.value),
               "preparation_seconds": 0.0, "solve_seconds": seconds.value, "batches": batches,
               "construction_steps": steps.value, "ls_evaluations": evals.value,
               "allocated_device_bytes": 0}
        if trace:
            out["trace"] = tr
        return out


# ---------------------------------------------------------------------------
# Diagnostics (parity harness; reference cuda/faco_operations.cu:60-132)

LOCAL_SEARCH_MODES = {"2opt": 0, "2opt+oropt": 1}


def _local_search_mode(mode) -> int:
    if mode not in LOCAL_SEARCH_MODES:
        raise ValueError("local search must be '2opt' or '2opt+oropt'")
    return LOCAL_SEARCH_MODES[mode]


def diagnostic_construct_ls(coordinates, ls_rows, parents, visits, mne_targets, evaluation_limit, waves_per_block=1,
                            ants_per_block=0, local_search="2opt"):
    xy = _coordinates(coordinates)
    n = xy.shape[0]
    ls_rows = np.ascontiguousarray(ls_rows, dtype=np.uint32)
    parents = np.ascontiguousarray(parents, dtype=np.uint32)
    visits = np.ascontiguousarray(visits, dtype=np.uint32)
    targets = np.ascontiguousarray(mne_targets, dtype=np.uint32)
    ants = parents.shape[0]
    if parents.shape != (ants, n) or visits.shape != (ants, n) or targets.shape != (ants,):
        raise ValueError("parents/visits must be [ants, n] and mne_targets [ants]")
    out = {k: np.zeros((ants, n), dtype=np.uint32) for k in ("construction_tours", "tours", "positions")}
    checklists = np.zeros((ants, 5 * n), dtype=np.uint32)
    info = np.zeros(ants, dtype=_native.ANT_INFO_DTYPE)
    check(lib().faco_diagnostic_construct_ls(xy.reshape(-1), n, ls_rows.reshape(-1), ls_rows.shape[1], ants,
                                             parents.reshape(-1), visits.reshape(-1), targets,
                                             _unsigned(evaluation_limit, "evaluation_limit"), waves_per_block,
                                             ants_per_block,
                                             out["construction_tours"].reshape(-1), out["tours"].reshape(-1),
                                             out["positions"].reshape(-1), checklists.reshape(-1), ptr(info),
                                             _local_search_mode(local_search)))
    out["checklists"] = checklists
    out["info"] = info
    return out


def diagnostic_distances(pairs) -> np.ndarray:
    pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 4)
    out = np.zeros(pairs.shape[0], dtype=np.int64)
    check(lib().faco_diagnostic_distances(pairs.reshape(-1), pairs.shape[0], out))
    return out


def diagnostic_philox(key, batch, ant, step, words=8) -> np.ndarray:
    out = np.zeros(words, dtype=np.uint32)
    check(lib().faco_diagnostic_philox(key, batch, ant, step, words, out))
    return out


def diagnostic_escape_rows(original, backup, n, from_cities, key, batch, events, ls=False, coordinates=None) -> dict:
    """Escape slot views of rows (original [rows][width <= 20], backup
    [rows][bw <= 64], padding = n) built on the device with the escape stream
    of (key, batch, row, events[row], ls); with coordinates, ordered as LS rows
    by (EUC_2D distance, id)."""
    o = np.ascontiguousarray(original, dtype=np.uint32)
    b = np.ascontiguousarray(backup, dtype=np.uint32)
    rows, width = o.shape
    bw = b.shape[1] if b.ndim == 2 else 0
    f = np.ascontiguousarray(from_cities, dtype=np.uint32)
    e = np.ascontiguousarray(events, dtype=np.uint32)
    if b.shape[0] != rows or f.size != rows or e.size != rows:
        raise ValueError("one backup row, city and event per row")
    xy = None if coordinates is None else _coordinates(coordinates)
    members = np.zeros((rows, width), dtype=np.uint32)
    replaced = np.zeros((rows, width), dtype=np.uint8)
    counts = np.zeros((rows, 2), dtype=np.uint32)
    check(lib().faco_diagnostic_escape_rows(rows, width, o.reshape(-1), bw, b.reshape(-1) if bw else np.zeros(1, np.uint32),
                                            n, f, _unsigned(key, "key"), _unsigned(batch, "batch", 32), e,
                                            1 if ls else 0, None if xy is None else xy.ctypes.data_as(ctypes.c_void_p),
                                            members.reshape(-1), replaced.reshape(-1), counts.reshape(-1)))
    return {"members": members, "replaced": replaced, "valid": counts[:, 0], "replacements": counts[:, 1]}


def diagnostic_escape_transactions(counts, removed, added, permits, cache_edges, cache_sizes, cache_epochs,
                                   offsets, neighbors) -> dict:
    """Escape edge transactions on the device: prepare each case's proposal
    (removed/added [cases][3][2]) against the CSR graph and its cache
    ([cases][64][2], size, epoch), then commit it."""
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    cases = c.size
    off = np.ascontiguousarray(offsets, dtype=np.uint32)
    nb = np.ascontiguousarray(neighbors, dtype=np.uint32)
    proposals = np.zeros((cases, 11), dtype=np.uint32)
    committed = np.zeros(cases, dtype=np.uint32)
    after = np.zeros((cases, 64, 2), dtype=np.uint32)
    size_after = np.zeros(cases, dtype=np.uint32)
    check(lib().faco_diagnostic_escape_transactions(
        cases, c, np.ascontiguousarray(removed, dtype=np.uint32).reshape(-1),
        np.ascontiguousarray(added, dtype=np.uint32).reshape(-1), np.ascontiguousarray(permits, dtype=np.uint32),
        np.ascontiguousarray(cache_edges, dtype=np.uint32).reshape(-1), np.ascontiguousarray(cache_sizes, dtype=np.uint32),
        np.ascontiguousarray(cache_epochs, dtype=np.uint64), off, nb if nb.size else np.zeros(1, np.uint32),
        off.size - 1, proposals.reshape(-1), committed, after.reshape(-1), size_after))
    return {"proposals": proposals, "committed": committed, "cache_after": after, "size_after": size_after}


def diagnostic_escape_draws(key, batch, ant, event, ls, count) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint32)
    check(lib().faco_diagnostic_escape_draws(_unsigned(key, "key"), batch, ant, event, 1 if ls else 0, count, out))
    return out


def diagnostic_copy(nbytes: int, repeats: int = 1) -> float:
    """Average milliseconds of the calibration copy kernel over `nbytes`."""
    ms = ctypes.c_double()
    check(lib().faco_diagnostic_copy(nbytes, repeats, ctypes.byref(ms)))
    return ms.value


def diagnostic_start_nodes(key: int, batch: int, ant: int, counts, on_device: bool = False) -> np.ndarray:
    """Start cities of the rejection sampler for each count (host or device)."""
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    out = np.zeros(counts.size, dtype=np.uint32)
    check(lib().faco_diagnostic_start_nodes(key, batch, ant, counts, counts.size, int(on_device), out))
    return out
