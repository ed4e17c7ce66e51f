"""Device discovery and a resident solver session on ROCm: the counterpart of
the reference's GpuSession and worker start-up checks
(python/gp_faco/gpu_session.py:17-176, python/gp_faco/worker.py:465-541),
which read nvidia-smi. Here the devices come from the amdgpu driver's own
sysfs files, read without creating a HIP context (so the device can still be
chosen through ROCR_VISIBLE_DEVICES before the engine library starts):

- KFD topology nodes (/sys/class/kfd/kfd/topology/nodes/N/properties): the
  GPU nodes (gpu_id != 0) with their unique id, gfx target version, SIMD and
  XCC counts and DRM render minor;
- the render node's device directory (/sys/class/drm/renderD<minor>/device):
  product name, VRAM total and used, busy percent (each optional: a file the
  driver does not provide reads as None);
- /sys/class/kfd/kfd/proc/<pid>/vram_<gpu_id>: the processes holding memory
  on that GPU, when the driver exposes them.

A session holds an exclusive lock file per GPU, refuses a device that is
not gfx950, is busy or is used by another process, and then keeps one
engine per panel shape (dimension, population size).
"""
from __future__ import annotations

import fcntl
import math
import os
import time
from pathlib import Path

import numpy as np

GFX950 = 90500   # gfx_target_version of gfx950 (major * 10000 + minor * 100 + stepping)


def faco_ants(dimension: int) -> int:
    """Ants per batch for n cities: 64 * ceil(sqrt(n) / 16) in integers
    (reference worker.py:30-34)."""
    if type(dimension) is not int or dimension < 3:
        raise ValueError("TSP dimension must be an integer of at least 3")
    return 64 * (math.isqrt(dimension - 1) // 16 + 1)


def _read(path: Path):
    try:
        return path.read_text().strip()
    except OSError:
        return None


def _integer(path: Path):
    text = _read(path)
    try:
        return int(text) if text is not None else None
    except ValueError:
        return None


def gfx_name(version: int) -> str:
    """gfx target name of a KFD gfx_target_version (90500 -> gfx950)."""
    major, minor, stepping = version // 10000, (version // 100) % 100, version % 100
    return f"gfx{major}{minor}{stepping:x}"


def list_gpus(sysfs: str | os.PathLike = "/sys") -> list[dict]:
    """Every GPU node of the KFD topology, in node order."""
    root = Path(sysfs)
    nodes = root / "class" / "kfd" / "kfd" / "topology" / "nodes"
    if not nodes.is_dir():
        return []
    processes = root / "class" / "kfd" / "kfd" / "proc"
    gpus = []
    for node in sorted(nodes.iterdir(), key=lambda p: int(p.name) if p.name.isdigit() else -1):
        gpu_id = _integer(node / "gpu_id")
        if not gpu_id:
            continue
        props = {}
        for line in (_read(node / "properties") or "").splitlines():
            parts = line.split()
            if len(parts) == 2 and parts[1].lstrip("-").isdigit():
                props[parts[0]] = int(parts[1])
        version = props.get("gfx_target_version", 0)
        simd_per_cu = props.get("simd_per_cu", 4) or 4
        minor = props.get("drm_render_minor")
        device = root / "class" / "drm" / f"renderD{minor}" / "device" if minor is not None else None
        users = []
        if processes.is_dir():
            for proc in processes.iterdir():
                if proc.name.isdigit() and (proc / f"vram_{gpu_id}").exists():
                    users.append(int(proc.name))
        gpus.append({
            "node": int(node.name), "gpu_id": gpu_id,
            "uuid": f"GPU-{props.get('unique_id', 0):016x}",
            "gfx_target_version": version, "gfx": gfx_name(version),
            "compute_units": props.get("simd_count", 0) // simd_per_cu, "xcc": props.get("num_xcc", 1),
            "render_minor": minor, "location_id": props.get("location_id"),
            "max_clock_mhz": props.get("max_engine_clk_fcompute"),
            "product": _read(device / "product_name") if device else None,
            "vram_total": _integer(device / "mem_info_vram_total") if device else None,
            "vram_used": _integer(device / "mem_info_vram_used") if device else None,
            "busy_percent": _integer(device / "gpu_busy_percent") if device else None,
            "processes": sorted(users),
        })
    return gpus


def check_idle(gpu: dict, *, own_pid: int | None = None) -> None:
    """The reference's start condition (an idle device of the expected
    model, gpu_session.py:20-45), on ROCm: gfx950, busy at most 5%, at most
    1 GiB of VRAM used and no other process on it (figures the driver does
    not report are not checked). Raises RuntimeError otherwise."""
    if gpu["gfx_target_version"] != GFX950:
        raise RuntimeError(f"a session needs an MI355X (gfx950), found {gpu['gfx']}")
    foreign = [p for p in gpu["processes"] if p != (own_pid if own_pid is not None else os.getpid())]
    if foreign:
        raise RuntimeError(f"GPU {gpu['uuid']} is used by processes {foreign}")
    if gpu["busy_percent"] is not None and gpu["busy_percent"] > 5:
        raise RuntimeError(f"GPU {gpu['uuid']} is busy ({gpu['busy_percent']}%)")
    if gpu["vram_used"] is not None and gpu["vram_used"] > 1 << 30:
        raise RuntimeError(f"GPU {gpu['uuid']} has {gpu['vram_used']} bytes of VRAM in use")


def decider_kind(controller) -> str:
    """Which evaluation a session panel runs for `controller`: "faco" (None:
    plain FACO with MNE 8, or {"uniform_mne": m}), "controller" (a
    controller dict), "program" (an IR program dict) or "baseline" (a
    baseline policy dict)."""
    if controller is None or "uniform_mne" in controller:
        return "faco"
    if "controller_version" in controller:
        return "controller"
    return "program" if "opcode" in controller else "baseline"


class _Panel:
    """One engine of a session and the problem names registered in it."""

    def __init__(self, engine):
        self.engine = engine
        self.names = set()

    def register(self, problems, names) -> float:
        """Registers each name not yet in the engine; returns the seconds spent."""
        started = time.perf_counter()
        for name in names:
            if name in self.names:
                continue
            problem = problems[name]
            self.engine.register_problem(problem.numeric_id, np.asarray(problem.coordinates, dtype=np.float64))
            self.names.add(name)
        return time.perf_counter() - started


class GpuSession:
    """Resident FACO solver on one GPU (reference gpu_session.py:17-176).

    gpu_uuid is the device's "GPU-<unique id>" (list_gpus); backend must be
    this engine's numeric backend ("euc2d_int32"). The GPU is chosen with
    ROCR_VISIBLE_DEVICES before the engine library starts; a process whose
    library already runs keeps its device, which must then be the requested
    one. Engines are kept per (dimension, population size) and reused by
    later panels of that shape; changing the colony count drops them."""

    def __init__(self, gpu_uuid, backend="euc2d_int32", *, colonies=32, lock_directory=None, sysfs="/sys",
                 idle_wait_seconds=0.0):
        _check_colonies(colonies)
        self.gpu_uuid, self.backend, self.colonies = gpu_uuid, backend, colonies
        directory = Path(lock_directory) if lock_directory else Path(os.environ.get("TMPDIR", "/tmp"))
        directory.mkdir(parents=True, exist_ok=True)
        self.lease = (directory / f"worker-{gpu_uuid}.lock").open("a")
        try:
            fcntl.flock(self.lease, fcntl.LOCK_EX | fcntl.LOCK_NB)
            gpu = wait_until_idle(gpu_uuid, sysfs=sysfs, wait_seconds=idle_wait_seconds)
            native, info = _load_engine_library(gpu, backend)
        except BaseException:
            self.lease.close()
            raise
        self.native = native
        driver = _read(Path(sysfs) / "module" / "amdgpu" / "version")
        self.device = {"gpu_uuid": gpu_uuid, "model": gpu["product"] or info.get("name"), "driver": driver,
                       "build": "libfaco_hip", "gfx": gpu["gfx"], "compute_units": gpu["compute_units"]}
        self._panels = {}

    def _panel(self, n, population=1) -> _Panel:
        shape = (n, population)
        if shape not in self._panels:
            settings = self.native.FixedFacoSettings()
            settings.ants = faco_ants(n)
            self._panels[shape] = _Panel(self.native.FacoBatchEngine(n, self.colonies, settings,
                                                                     population_size=population))
        return self._panels[shape]

    def _tasks(self, problems, pairs):
        """Dimension, instance keys and seeds of a complete panel of one size."""
        sizes = {problems[name].dimension for name, _ in pairs}
        if len(sizes) != 1:
            raise ValueError("a panel cannot mix dimensions")
        keys = np.fromiter((problems[name].numeric_id for name, _ in pairs), dtype=np.uint64, count=len(pairs))
        seeds = np.fromiter((seed for _, seed in pairs), dtype=np.uint64, count=len(pairs))
        return sizes.pop(), keys, seeds

    def solve(self, problems, pairs, iterations, controller, *, checkpoints=(), decision_iterations=()):
        """One panel: pairs = [(problem name, seed)] * colonies; controller
        None (plain FACO, MNE 8), {"uniform_mne": m}, a controller, a
        program or a baseline policy dict (reference gpu_session.py:62-129).
        Returns the engine's result and the registration / solve seconds and
        device bytes."""
        if len(pairs) != self.colonies:
            raise ValueError("FACO needs one complete fixed panel")
        n, keys, seeds = self._tasks(problems, pairs)
        panel = self._panel(n)
        registration_seconds = panel.register(problems, [name for name, _ in pairs])
        budget = faco_ants(n) * iterations
        kind = decider_kind(controller)
        extra = {"decision_iterations": list(decision_iterations)} if decision_iterations else {}
        if kind != "controller":
            extra["checkpoint_iterations"] = list(checkpoints)
        engine = panel.engine
        started = time.perf_counter()
        if kind == "faco":
            mne = 8 if controller is None else controller["uniform_mne"]
            result = engine.evaluate_faco_evaluations(keys, seeds, budget, mne, "cached", **extra)
        elif kind == "controller":
            result = engine.evaluate_controller_evaluations(keys, seeds, budget, controller, **extra)
        elif kind == "program":
            result = engine.evaluate_program_evaluations(keys, seeds, budget, controller, "cached", **extra)
        else:
            result = engine.evaluate_baseline_evaluations(keys, seeds, budget, controller, "cached", **extra)
        return result, {"registration_seconds": registration_seconds, "solve_seconds": time.perf_counter() - started,
                        "device_bytes": result["allocated_device_bytes"]}

    def solve_population(self, problems, pairs, iterations, programs, *, experiment_mask=0xFFFFFFFF):
        """The panel under 1..128 programs in one engine (reference
        gpu_session.py:139-171): colony d * panel + t is task t under program d."""
        if len(pairs) != self.colonies or not 1 <= len(programs) <= 128:
            raise ValueError("a population shard needs a complete panel and 1..128 programs")
        n, keys, seeds = self._tasks(problems, pairs)
        started = time.perf_counter()
        panel = self._panel(n, len(programs))
        panel.register(problems, [name for name, _ in pairs])
        registration_seconds = time.perf_counter() - started
        started = time.perf_counter()
        result = panel.engine.evaluate_population_evaluations(keys, seeds, faco_ants(n) * iterations, programs,
                                                              experiment_mask)
        return result, {"registration_seconds": registration_seconds, "solve_seconds": time.perf_counter() - started}

    def set_colonies(self, colonies):
        """A new panel size: the engines of the old size are dropped."""
        _check_colonies(colonies)
        if colonies != self.colonies:
            self._panels = {}
            self.colonies = colonies

    def close(self):
        self._panels = {}
        self.lease.close()


def _check_colonies(colonies):
    if not 1 <= colonies <= 128:
        raise ValueError("the GPU panel must hold 1..128 colonies")


def wait_until_idle(gpu_uuid, *, sysfs="/sys", wait_seconds=0.0, poll_seconds=0.2) -> dict:
    """The device record of gpu_uuid once check_idle passes, polling for up to
    wait_seconds (the reference worker waits for idle readings while holding
    its lock, worker.py:465-494). A device that is not gfx950 fails at once."""
    deadline = time.monotonic() + wait_seconds
    while True:
        matches = [g for g in list_gpus(sysfs) if g["uuid"] == gpu_uuid]
        if len(matches) != 1:
            raise RuntimeError(f"no GPU {gpu_uuid} in the KFD topology")
        gpu = matches[0]
        try:
            check_idle(gpu)
            return gpu
        except RuntimeError:
            if gpu["gfx_target_version"] != GFX950 or time.monotonic() >= deadline:
                raise
        time.sleep(poll_seconds)


def _load_engine_library(gpu, backend):
    """Binds the process to `gpu` (ROCR_VISIBLE_DEVICES, when the engine
    library has not started yet) and loads it; checks the numeric backend and
    that the visible HIP device is that GPU."""
    from . import _native
    if _native._lib is None:
        os.environ["ROCR_VISIBLE_DEVICES"] = gpu["uuid"]
    from . import gp_faco_ext as native
    if native.numeric_backend != backend:
        raise ValueError("one process loads one numeric backend")
    info = native.cuda_device_info()
    if info.get("arch", "").split(":")[0] != gpu["gfx"]:
        raise RuntimeError("the visible HIP device is not the requested GPU")
    return native, info
