"""The ROCm session on the real device: discovery finds the MI355X this
process runs on, and a session's panel solves equal the engine's own."""
from dataclasses import dataclass

import numpy as np
import pytest

from gplsaco_amd import gpu_session as G
from gplsaco_amd.instances import uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu


@dataclass
class Problem:
    numeric_id: int
    coordinates: np.ndarray

    @property
    def dimension(self):
        return self.coordinates.shape[0]


def test_discovery_and_session_solves(ext, tmp_path):
    gpus = G.list_gpus()
    announce("session", [(g["uuid"], g["gfx"]) for g in gpus])
    mine = [g for g in gpus if g["gfx"] == "gfx950"]
    assert mine, gpus
    g = mine[0]
    info = ext.cuda_device_info()
    assert g["compute_units"] == info["compute_units"] and g["xcc"] == 8 and g["vram_total"] > 250e9
    session = G.GpuSession(g["uuid"], lock_directory=tmp_path, colonies=4, idle_wait_seconds=30)
    try:
        assert session.device["gfx"] == "gfx950" and session.device["compute_units"] == 256
        problems = {f"p{i}": Problem(40 + i, uniform_instance(300, seed=12, instance=i)) for i in range(2)}
        pairs = [("p0", 1), ("p1", 1), ("p0", 2), ("p1", 2)]
        result, timing = session.solve(problems, pairs, 5, None)
        engine = ext.FacoBatchEngine(300, 4, ext.FixedFacoSettings())
        for p in problems.values():
            engine.register_problem(p.numeric_id, p.coordinates)
        keys = np.array([40, 41, 40, 41], dtype=np.uint64)
        seeds = np.array([1, 1, 2, 2], dtype=np.uint64)
        want = engine.evaluate_faco_evaluations(keys, seeds, G.faco_ants(300) * 5)
        assert [i["tour"] for i in result["items"]] == [i["tour"] for i in want["items"]]
        assert timing["device_bytes"] > 0 and timing["solve_seconds"] > 0
        policy = {"policy_spec_id": 1, "kind": "static", "mne_level": 1, "max_mne_level": 1, "region": 2,
                  "restart_mode": "none", "restart_period": 0, "restart_probability": 0.0, "stagnation_step": 0,
                  "restart_stagnation": 0, "restart_cooldown": 0}
        base, _ = session.solve(problems, pairs, 5, policy, decision_iterations=[5])
        want = engine.evaluate_baseline_evaluations(keys, seeds, G.faco_ants(300) * 5, policy, decision_iterations=[5])
        assert [i["tour"] for i in base["items"]] == [i["tour"] for i in want["items"]]
        assert base["decisions"] == want["decisions"]
        with pytest.raises(BlockingIOError):
            G.GpuSession(g["uuid"], lock_directory=tmp_path)
    finally:
        session.close()
