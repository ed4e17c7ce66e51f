"""A resident single-GPU worker process on ROCm: the counterpart of the
reference's WorkerProtocol and PersistentGpuWorker
(python/gp_faco/worker.py:57-200, 453-877) for population evaluations.

The coordinating process validates the protocol and the tasks and never
loads the engine library; one spawned worker process does, after its device
checks (worker.py:465-541, here on the amdgpu driver's sysfs files instead
of nvidia-smi):

1. an exclusive lock file per GPU (one worker per device);
2. the device named by the protocol is in the KFD topology, is gfx950, is
   the protocol's model, and is idle: no other process holds memory on it,
   at most 5% busy and at most 1 GiB of VRAM in use, polled for up to
   `idle_wait_seconds` (processes listed in `shared_with` may hold it too:
   a harness that already uses the GPU);
3. ROCR_VISIBLE_DEVICES binds the process to that GPU before the library
   starts, and the visible HIP device and numeric backend ("euc2d_int32")
   are checked.

Tasks are evaluation-count GP programs over a fixed panel (colonies x
instances x seeds), as the reference's population entry takes them
(worker.py:707-814): 1..128 programs sharing instances, replicas, budget and
mask run as one population evaluation, each member's result the native
result dict of evaluate_population_evaluations.
"""
from __future__ import annotations

import math
import multiprocessing
import os
import platform
import re
import time
from concurrent.futures import Future, ProcessPoolExecutor
from dataclasses import asdict, dataclass, field
from pathlib import Path

import numpy as np

from . import gpu_session

PROJECT = Path(__file__).resolve().parents[1]
BACKEND = "euc2d_int32"


@dataclass(frozen=True)
class SolverSettings:
    """FACO settings of a worker (reference worker.py:57-100); ants 0 = the
    per-size default."""
    ants: int = 0
    primary_width: int = 16
    backup_width: int = 64
    ls_width: int = 20
    beta: float = 1.0
    retention: float = 0.5
    p_best: float = 0.1
    epoch_source_probability: float = 0.01
    ls_evaluation_limit: int = 100000
    initial_ls_evaluation_limit: int = 100000

    def __post_init__(self) -> None:
        for name in ("ants", "primary_width", "backup_width", "ls_width", "ls_evaluation_limit",
                     "initial_ls_evaluation_limit"):
            value = getattr(self, name)
            if type(value) is not int or not 0 <= value <= 0xFFFFFFFFFFFFFFFF:
                raise ValueError(f"{name} must be a non-negative integer")
        if not 0 <= self.ants <= 4096 or self.primary_width < 2 or self.ls_width < 1:
            raise ValueError("invalid ant count or candidate widths")
        if any(getattr(self, name) > 0xFFFFFFFF for name in ("primary_width", "backup_width", "ls_width")):
            raise ValueError("candidate width beyond uint32")
        for name in ("beta", "retention", "p_best", "epoch_source_probability"):
            value = getattr(self, name)
            if type(value) not in (int, float) or not math.isfinite(value):
                raise ValueError(f"{name} must be finite")
        if not (self.beta > 0 and 0 <= self.retention < 1 and 0 < self.p_best < 1
                and 0 <= self.epoch_source_probability <= 1):
            raise ValueError("pheromone parameters out of range")

    def ants_for(self, dimension: int) -> int:
        return self.ants or gpu_session.faco_ants(dimension)


@dataclass(frozen=True)
class Problem:
    """An unlabelled instance: a stable positive numeric key (the engine's
    instance key), a name and its coordinates."""
    instance_id: str
    numeric_id: int
    coordinates: tuple

    def __post_init__(self) -> None:
        if type(self.instance_id) is not str or not self.instance_id:
            raise ValueError("an instance needs a non-empty name")
        if type(self.numeric_id) is not int or not 0 < self.numeric_id < 1 << 64:
            raise ValueError("an instance needs a stable positive numeric key")
        xy = np.asarray(self.coordinates, dtype=np.float64)
        if xy.ndim != 2 or xy.shape[1] != 2 or xy.shape[0] < 3:
            raise ValueError("coordinates must be [n][2] with n >= 3")
        object.__setattr__(self, "coordinates", tuple(map(tuple, xy.tolist())))

    @property
    def dimension(self) -> int:
        return len(self.coordinates)


def freeze_problems(values) -> tuple[Problem, ...]:
    """A panel: distinct names, one dimension, sorted by name (reference
    worker.py:41-54)."""
    problems = sorted(values, key=lambda p: p.instance_id)
    if any(type(p) is not Problem for p in problems):
        raise TypeError("worker problems must be Problem records")
    names = [p.instance_id for p in problems]
    if not names or len(set(names)) != len(names) or len({p.dimension for p in problems}) != 1:
        raise ValueError("a panel needs distinct instances of one dimension")
    return tuple(problems)


@dataclass(frozen=True)
class WorkerProtocol:
    """What a worker runs on and with (reference worker.py:102-200): the GPU
    (KFD unique id as "GPU-<16 hex digits>", product name, gfx target), the
    numeric backend, the panel shape and the solver settings. Graph-constrained
    workers need the reference's graph catalog, which this library does not
    carry; they are refused."""
    gpu_uuid: str
    gpu_model: str
    gfx: str = "gfx950"
    numeric_backend: str = BACKEND
    build_id: str = "mi355x-hip"
    dimensions: tuple = (500, 1000)
    colonies: int = 32
    settings: SolverSettings = field(default_factory=SolverSettings)
    maximum_registered_per_dimension: int = 1024
    execution_host: str = field(default_factory=platform.node)
    constraint_mode: str = "unrestricted"
    lock_directory: str | None = None
    sysfs: str = "/sys"
    idle_wait_seconds: float = 0.0
    shared_with: tuple = ()

    def __post_init__(self) -> None:
        object.__setattr__(self, "dimensions", tuple(self.dimensions))
        object.__setattr__(self, "shared_with", tuple(self.shared_with))
        if type(self.gpu_uuid) is not str or not re.fullmatch(r"GPU-[0-9a-f]{16}", self.gpu_uuid):
            raise ValueError("a worker needs the GPU's full KFD unique id (GPU-<16 hex digits>)")
        if not self.gpu_model:
            raise ValueError("the GPU model must be named explicitly")
        if self.gfx != "gfx950":
            raise ValueError("this worker runs on MI355X (gfx950) only")
        if self.numeric_backend != BACKEND:
            raise ValueError("unknown numeric backend")
        if (not self.dimensions or len(set(self.dimensions)) != len(self.dimensions)
                or any(type(n) is not int or not 3 <= n <= 10000 for n in self.dimensions)):
            raise ValueError("dimensions must be distinct integers in 3..10000")
        if type(self.colonies) is not int or not 1 <= self.colonies <= 128:
            raise ValueError("the panel must hold 1..128 colonies")
        if type(self.maximum_registered_per_dimension) is not int or self.maximum_registered_per_dimension < self.colonies:
            raise ValueError("the instance cache must hold a whole panel")
        if type(self.settings) is not SolverSettings:
            raise TypeError("settings must be SolverSettings")
        if self.constraint_mode not in ("unrestricted", "hard", "escape"):
            raise ValueError("unknown constraint mode")
        if self.constraint_mode != "unrestricted":
            raise ValueError("graph-constrained workers need the graph catalog, which this library does not carry")
        if type(self.idle_wait_seconds) not in (int, float) or not 0 <= self.idle_wait_seconds < math.inf:
            raise ValueError("idle wait must be a finite non-negative number of seconds")
        if any(type(pid) is not int or pid <= 0 for pid in self.shared_with):
            raise ValueError("shared_with lists process ids")

    def manifest(self) -> dict:
        out = asdict(self)
        out.update({"protocol_version": 6, "local_search_spec": "checklist_2opt_identity_zero_v2",
                    "instance_key_policy": "dataset_record_ordinal_v2",
                    "submission_boundary": "whole_panel_returned_final_tours_scored_once",
                    "gpu_lease_policy": "lock_file_per_kfd_unique_id"})
        return out

    @property
    def identifier(self) -> str:
        return f"{self.build_id}-{self.constraint_mode}-knn"


@dataclass(frozen=True)
class SolveTask:
    """One GP program over a panel for an evaluation count (reference
    worker.py:251-342, count form): `replicas` lists (instance name, seed)
    per colony."""
    occurrence_id: str
    program: dict
    problems: tuple
    replicas: tuple
    evaluation_limit_per_colony: int
    experiment_mask: int = 0xFFFFFFFF
    preparation_mode: str = "cached"

    def __post_init__(self) -> None:
        from .gp_faco_ext import read_program   # pure-Python IR check; loads no library
        if type(self.occurrence_id) is not str or not self.occurrence_id:
            raise ValueError("every evaluated position needs its own identity")
        read_program(self.program)
        if self.program["feature_spec_id"] != 2:
            raise ValueError("evaluation-count tasks read feature spec 2")
        object.__setattr__(self, "problems", freeze_problems(self.problems))
        object.__setattr__(self, "replicas", tuple(tuple(row) for row in self.replicas))
        names = {p.instance_id for p in self.problems}
        if any(len(row) != 2 or type(row[0]) is not str or type(row[1]) is not int
               or not 0 <= row[1] < 1 << 64 for row in self.replicas):
            raise ValueError("a replica is an instance name and a uint64 seed")
        if len(set(self.replicas)) != len(self.replicas) or {row[0] for row in self.replicas} != names:
            raise ValueError("replicas miss an instance, name an unknown one or repeat one")
        if type(self.evaluation_limit_per_colony) is not int or not 0 <= self.evaluation_limit_per_colony <= 4096 * 0xFFFFFFFF:
            raise ValueError("an evaluation-count task needs an integer FE budget")
        if self.preparation_mode not in ("cached", "end_to_end"):
            raise ValueError("unknown preparation mode")
        if type(self.experiment_mask) is not int or not 0 < self.experiment_mask <= 0xFFFFFFFF or not self.experiment_mask & 0xFFFF:
            raise ValueError("the mask must keep at least one hold action")

    @property
    def dimension(self) -> int:
        return self.problems[0].dimension

    def manifest(self, protocol: WorkerProtocol) -> dict:
        ants = protocol.settings.ants_for(self.dimension)
        if self.dimension not in protocol.dimensions or len(self.replicas) != protocol.colonies:
            raise ValueError("the task does not fit the worker's dimensions or panel shape")
        if self.evaluation_limit_per_colony % ants or self.evaluation_limit_per_colony // ants > 0xFFFFFFFF:
            raise ValueError("the budget must be a whole number of ant batches")
        return {"occurrence_id": self.occurrence_id, "protocol_id": protocol.identifier,
                "dimension": self.dimension, "problems": [p.instance_id for p in self.problems],
                "replicas": self.replicas, "evaluation_limit_per_colony": self.evaluation_limit_per_colony,
                "preparation_mode": self.preparation_mode, "experiment_mask": self.experiment_mask}


def population_members(tasks, protocol: WorkerProtocol) -> SolveTask:
    """Checks a population batch (reference worker.py:707-734): 1..128 tasks
    sharing instances, replicas, budget, mask and preparation ("cached");
    returns the first."""
    if not tasks or len(tasks) > 128:
        raise ValueError("a population batch holds 1..128 individuals")
    first = tasks[0]

    def signature(t):
        return (type(t), t.dimension, t.replicas, t.evaluation_limit_per_colony, t.preparation_mode,
                t.experiment_mask, tuple((p.instance_id, p.numeric_id) for p in t.problems))
    for task in tasks:
        if type(task) is not SolveTask or signature(task) != signature(first):
            raise ValueError("a population shares instances, seeds, budget and mask")
        if task.preparation_mode != "cached":
            raise ValueError("a population takes the cached count entry only")
        task.manifest(protocol)
    return first


# ---------------------------------------------------------------------------
# The worker process's side: module state and entry points (run through the
# spawned executor).
_state: dict = {}


def _initialize(protocol: WorkerProtocol) -> dict:
    import fcntl
    started = time.perf_counter()
    if protocol.execution_host != platform.node():
        raise RuntimeError("the execution host differs from the protocol's")
    directory = Path(protocol.lock_directory) if protocol.lock_directory else PROJECT / ".tmp"
    directory.mkdir(parents=True, exist_ok=True)
    lease = (directory / f"worker-{protocol.gpu_uuid}.lock").open("a")
    try:
        fcntl.flock(lease, fcntl.LOCK_EX | fcntl.LOCK_NB)
    except BlockingIOError:
        lease.close()
        raise BlockingIOError(f"another worker holds {protocol.gpu_uuid}") from None
    gpu = _wait_for_idle(protocol)
    if gpu["product"] is not None and gpu["product"] != protocol.gpu_model:
        raise RuntimeError(f"GPU {protocol.gpu_uuid} is a {gpu['product']}, not a {protocol.gpu_model}")
    native, info = gpu_session._load_engine_library(gpu, protocol.numeric_backend)
    _state.update(protocol=protocol, lease=lease, native=native, population={})
    return {"pid": os.getpid(), "host": platform.node(), "gpu": gpu, "device": info,
            "library": str(Path(native.__file__).resolve().parent / "libfaco_hip.so"),
            "initialization_seconds": time.perf_counter() - started}


def _wait_for_idle(protocol: WorkerProtocol) -> dict:
    """The device record once it is idle (reference worker.py:465-494): no
    other process on it (besides `shared_with`), busy <= 5%, VRAM used <=
    1 GiB, polled for up to idle_wait_seconds."""
    deadline = time.monotonic() + protocol.idle_wait_seconds
    while True:
        matches = [g for g in gpu_session.list_gpus(protocol.sysfs) if g["uuid"] == protocol.gpu_uuid]
        if len(matches) != 1:
            raise RuntimeError(f"no GPU {protocol.gpu_uuid} in the KFD topology")
        gpu = matches[0]
        if gpu["gfx"] != protocol.gfx:
            raise RuntimeError(f"GPU {protocol.gpu_uuid} is {gpu['gfx']}, not {protocol.gfx}")
        shared = set(protocol.shared_with) | {os.getpid()}
        foreign = [p for p in gpu["processes"] if p not in shared]
        if foreign:
            raise BlockingIOError(f"GPU {protocol.gpu_uuid} is used by processes {foreign}")
        busy = gpu["busy_percent"] is not None and gpu["busy_percent"] > 5
        full = not protocol.shared_with and gpu["vram_used"] is not None and gpu["vram_used"] > 1 << 30
        if not busy and not full:
            return gpu
        if time.monotonic() >= deadline:
            raise RuntimeError(f"GPU {protocol.gpu_uuid} did not become idle "
                               f"(busy {gpu['busy_percent']}%, VRAM used {gpu['vram_used']})")
        time.sleep(0.2)


def _initialize_and_ready(protocol: WorkerProtocol) -> dict:
    return _initialize(protocol)


def _execute_population(tasks) -> dict:
    """One population evaluation (reference worker.py:737-814): a cached
    population engine per dimension (replaced when too small or full),
    instances registered once, every member's native result."""
    started = time.perf_counter()
    protocol, native = _state["protocol"], _state["native"]
    first = population_members(tasks, protocol)
    n, count = first.dimension, len(tasks)
    outputs = [{"task_id": t.occurrence_id, "occurrence_id": t.occurrence_id, "protocol_id": protocol.identifier,
                "dimension": n, "worker_pid": os.getpid()} for t in tasks]
    try:
        cached = _state["population"].get(n)
        incoming = {p.numeric_id for p in first.problems}
        if cached is not None and (cached["count"] < count or
                                   len(cached["registered"] | incoming) > protocol.maximum_registered_per_dimension):
            del _state["population"][n]
            cached = None
        if cached is None:
            s = native.FixedFacoSettings()
            for name, value in asdict(protocol.settings).items():
                setattr(s, name, value)
            cached = {"count": count, "registered": set(), "fees": {},
                      "engine": native.FacoBatchEngine(n, protocol.colonies, s, population_size=count)}
            _state["population"][n] = cached
        engine = cached["engine"]
        for problem in first.problems:
            if problem.numeric_id not in cached["registered"]:
                cached["fees"][problem.instance_id] = engine.register_problem(
                    problem.numeric_id, np.asarray(problem.coordinates, dtype=np.float64))
                cached["registered"].add(problem.numeric_id)
        key_of = {p.instance_id: p.numeric_id for p in first.problems}
        keys = np.asarray([key_of[name] for name, _ in first.replicas], dtype=np.uint64)
        seeds = np.asarray([seed for _, seed in first.replicas], dtype=np.uint64)
        results = engine.evaluate_population_evaluations(keys, seeds, first.evaluation_limit_per_colony,
                                                         [t.program for t in tasks], first.experiment_mask)
        for output, result in zip(outputs, results, strict=True):
            output.update(status="completed", native_result=result,
                          registered_problems=len(cached["registered"]),
                          registration_fees={p.instance_id: cached["fees"][p.instance_id] for p in first.problems})
    except Exception as error:   # the whole batch fails together; a crash propagates through the Future
        for output in outputs:
            output.update(status="failed", error=f"{type(error).__name__}: {error}")
    elapsed = time.perf_counter() - started
    for output in outputs:
        output["worker_seconds"] = elapsed / count
        output["population_worker_seconds"] = elapsed
    return {"members": outputs}


class PersistentGpuWorker:
    """One spawned worker process per GPU with one active request at a time
    (reference worker.py:817-877); a Future timing out neither restarts the
    process nor resubmits."""

    def __init__(self, protocol: WorkerProtocol):
        if type(protocol) is not WorkerProtocol:
            raise TypeError("protocol must be a WorkerProtocol")
        self.protocol = protocol
        self._executor = ProcessPoolExecutor(max_workers=1, mp_context=multiprocessing.get_context("spawn"))
        self._ready_future = self._executor.submit(_initialize_and_ready, protocol)
        self._active: Future | None = None
        self._closed = False

    def ready(self, timeout: float | None = None) -> dict:
        return self._ready_future.result(timeout=timeout)

    def submit_population(self, tasks) -> Future:
        tasks = tuple(tasks)
        population_members(tasks, self.protocol)
        self._check_idle()
        self._active = self._executor.submit(_execute_population, tasks)
        return self._active

    def _check_idle(self) -> None:
        if self._closed:
            raise RuntimeError("the worker is closed")
        if self._active is not None and not self._active.done():
            raise RuntimeError("the worker already runs a request; wait for its Future")

    def close(self) -> None:
        if not self._closed:
            self._executor.shutdown(wait=True, cancel_futures=False)
            self._closed = True

    def __enter__(self) -> "PersistentGpuWorker":
        return self

    def __exit__(self, *args) -> None:
        self.close()
