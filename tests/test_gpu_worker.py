"""A population submitted through the ROCm worker (gplsaco_amd.worker): the
spawned worker passes its device checks on this box's MI355X, binds it, and
every member's result equals the same program's evaluation in this process
(reference python/gp_faco/worker.py:737-877)."""
import os

import numpy as np
import pytest

from gplsaco_amd import gpu_session, worker as W
from gplsaco_amd.instances import uniform_instance
from gplsaco_amd.programs import random_program
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def test_population_through_the_worker(ext, tmp_path):
    gpus = [g for g in gpu_session.list_gpus() if g["gfx"] == "gfx950"]
    assert gpus, "no gfx950 device in the KFD topology"
    gpu = gpus[0]
    n, colonies = 120, 4
    problems = tuple(W.Problem(f"inst{i}", 900 + i, uniform_instance(n, seed=17, instance=i)) for i in range(2))
    replicas = (("inst0", 3), ("inst0", 4), ("inst1", 3), ("inst1", 4))
    rng = np.random.default_rng(21)
    programs = [{**random_program(rng, depth=4), "feature_spec_id": 2} for _ in range(3)]
    ants = gpu_session.faco_ants(n)
    tasks = [W.SolveTask(f"member{i}", p, problems, replicas, ants * 10) for i, p in enumerate(programs)]
    # this test process may hold the GPU already (the library is loaded): it shares the device
    protocol = W.WorkerProtocol(gpu_uuid=gpu["uuid"], gpu_model=gpu["product"] or "unknown", dimensions=(n,),
                                colonies=colonies, lock_directory=str(tmp_path), idle_wait_seconds=30.0,
                                shared_with=(os.getpid(),))
    announce("worker-population", gpu["uuid"])
    with W.PersistentGpuWorker(protocol) as worker:
        receipt = worker.ready(timeout=300)
        assert receipt["gpu"]["uuid"] == gpu["uuid"] and receipt["pid"] != os.getpid()
        assert receipt["device"]["arch"].split(":")[0] == "gfx950"
        out = worker.submit_population(tasks).result(timeout=300)
        with W.PersistentGpuWorker(protocol) as second, pytest.raises(BlockingIOError):
            second.ready(timeout=300)                             # one worker per GPU: the lock is held
    members = out["members"]
    assert [m["status"] for m in members] == ["completed"] * 3, members[0].get("error")
    engine = ext.FacoBatchEngine(n, colonies, ext.FixedFacoSettings())
    for p in problems:
        engine.register_problem(p.numeric_id, np.asarray(p.coordinates))
    keys = np.array([900 + int(name[-1]) for name, _ in replicas], dtype=np.uint64)
    seeds = np.array([seed for _, seed in replicas], dtype=np.uint64)
    for member, program in zip(members, programs):
        want = engine.evaluate_program_evaluations(keys, seeds, ants * 10, program)
        got = member["native_result"]
        assert [(i["tour"], i["cost"]) for i in got["items"]] == [(i["tour"], i["cost"]) for i in want["items"]]
        assert got["control_states"] == want["control_states"]
