"""The ROCm worker's protocol, task rules and device checks on the host
(reference python/gp_faco/worker.py:57-200, 465-541, 707-734), over the
fake amdgpu sysfs tree of tests/test_gpu_session_host.py. No engine library
is loaded: every case here fails before the worker would bind a device."""
import fcntl
import os

import numpy as np
import pytest

from gplsaco_amd import worker as W
from gplsaco_amd.programs import random_program
from tests.test_gpu_session_host import UNIQUE, fake_sysfs

UUID = f"GPU-{UNIQUE:016x}"
MODEL = "AMD Instinct MI355 OAM"


def protocol(tmp_path, **kw):
    args = dict(gpu_uuid=UUID, gpu_model=MODEL, dimensions=(30,), colonies=4, lock_directory=str(tmp_path / "locks"),
                sysfs=str(tmp_path))
    args.update(kw)
    return W.WorkerProtocol(**args)


def problems(n=30, count=2):
    rng = np.random.default_rng(n)
    return tuple(W.Problem(f"p{i}", 100 + i, rng.integers(0, 1000, size=(n, 2))) for i in range(count))


def task(name="t0", limit=64 * 3, seed=0, **kw):
    program = {**random_program(np.random.default_rng(seed), depth=3), "feature_spec_id": 2}
    probs = problems()
    return W.SolveTask(name, program, probs, (("p0", 1), ("p0", 2), ("p1", 1), ("p1", 2)), limit, **kw)


def test_protocol_rules(tmp_path):
    p = protocol(tmp_path)
    assert p.identifier == "mi355x-hip-unrestricted-knn"
    m = p.manifest()
    assert m["gpu_uuid"] == UUID and m["protocol_version"] == 6 and m["numeric_backend"] == "euc2d_int32"
    for bad in (dict(gpu_uuid="GPU-1234"), dict(gpu_uuid="GPU-" + "0" * 36), dict(gpu_model=""),
                dict(gfx="gfx942"), dict(numeric_backend="fp32"), dict(dimensions=(30, 30)),
                dict(dimensions=(2,)), dict(dimensions=(30.0,)), dict(colonies=0), dict(colonies=129),
                dict(maximum_registered_per_dimension=3), dict(constraint_mode="hard"),
                dict(constraint_mode="escape"), dict(constraint_mode="other"), dict(idle_wait_seconds=-1),
                dict(shared_with=(0,))):
        with pytest.raises(ValueError):
            protocol(tmp_path, **bad)
    with pytest.raises(TypeError):
        protocol(tmp_path, settings={"ants": 0})
    with pytest.raises(ValueError):
        W.SolverSettings(retention=1.0)


def test_task_and_population_rules(tmp_path):
    p = protocol(tmp_path)
    t = task()
    assert t.dimension == 30 and [x.instance_id for x in t.problems] == ["p0", "p1"]
    assert t.manifest(p)["evaluation_limit_per_colony"] == 192
    with pytest.raises(ValueError):
        task(limit=100).manifest(p)                     # not a whole number of 64-ant batches
    with pytest.raises(ValueError):   # a feature-spec-1 (time-budget) program
        W.SolveTask("x", random_program(np.random.default_rng(0), depth=3), problems(), (("p0", 1), ("p1", 1)), 64)
    with pytest.raises(ValueError):
        W.SolveTask("x", task().program, problems(), (("p0", 1), ("p0", 1), ("p1", 1)), 64)   # repeated replica
    with pytest.raises(ValueError):
        W.SolveTask("x", task().program, problems(), (("p0", 1),), 64)                           # instance missing
    with pytest.raises(ValueError):
        task(experiment_mask=0xFFFF0000)
    tasks = [task(f"t{i}", seed=i) for i in range(3)]
    assert W.population_members(tasks, p) is tasks[0]
    with pytest.raises(ValueError):
        W.population_members(tasks + [task("odd", limit=128)], p)
    with pytest.raises(ValueError):
        W.population_members([task("e", preparation_mode="end_to_end")], p)
    with pytest.raises(ValueError):
        W.population_members([], p)
    with pytest.raises(ValueError):
        W.population_members([task()], protocol(tmp_path, colonies=8))   # panel shape differs


def test_device_checks(tmp_path):
    fake_sysfs(tmp_path)
    gpu = W._wait_for_idle(protocol(tmp_path))
    assert gpu["uuid"] == UUID and gpu["gfx"] == "gfx950" and gpu["product"] == MODEL
    with pytest.raises(RuntimeError):
        W._wait_for_idle(protocol(tmp_path, gpu_uuid="GPU-" + "0" * 16))
    other = tmp_path / "other"
    fake_sysfs(other, gfx=90402)
    with pytest.raises(RuntimeError):
        W._wait_for_idle(protocol(tmp_path, sysfs=str(other)))
    used = tmp_path / "used"
    fake_sysfs(used, users=(4242,))
    with pytest.raises(BlockingIOError):
        W._wait_for_idle(protocol(tmp_path, sysfs=str(used)))
    assert W._wait_for_idle(protocol(tmp_path, sysfs=str(used), shared_with=(4242,)))["processes"] == [4242]
    busy = tmp_path / "busy"
    fake_sysfs(busy, busy=40)
    with pytest.raises(RuntimeError):
        W._wait_for_idle(protocol(tmp_path, sysfs=str(busy), idle_wait_seconds=0.3))
    full = tmp_path / "full"
    fake_sysfs(full, used=2 << 30)
    with pytest.raises(RuntimeError):
        W._wait_for_idle(protocol(tmp_path, sysfs=str(full)))


def test_one_worker_per_gpu_and_checks_before_binding(tmp_path):
    fake_sysfs(tmp_path)
    p = protocol(tmp_path)
    locks = tmp_path / "locks"
    locks.mkdir()
    with open(locks / f"worker-{UUID}.lock", "a") as held:
        fcntl.flock(held, fcntl.LOCK_EX | fcntl.LOCK_NB)
        with pytest.raises(BlockingIOError):
            W._initialize(p)
    other = tmp_path / "other"
    fake_sysfs(other, gfx=90402)
    before = os.environ.get("ROCR_VISIBLE_DEVICES")
    with pytest.raises(RuntimeError):
        W._initialize(protocol(tmp_path, sysfs=str(other), lock_directory=str(tmp_path / "l2")))
    assert os.environ.get("ROCR_VISIBLE_DEVICES") == before      # refused before binding a device
    with pytest.raises(RuntimeError):
        W._initialize(protocol(tmp_path, execution_host="elsewhere"))
