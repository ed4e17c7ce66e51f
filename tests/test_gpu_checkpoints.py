"""Curve checkpoints and production-event timing on the device.

The reference's own check (tests/python/test_v2_cuda.py:17-36): the tours
recorded at iterations [0, 1, 3] of a 3-batch run are exact prefixes (each
equals a separate run of that many batches), costs never rise, the last one
is the final result, and unordered, repeated or out-of-budget points raise
ValueError. Here the prefixes are also pinned to the oracle's FACO run of
that many iterations, across graph replays (segments of 16 batches and eager
rests), for mixed-size batches (a finished colony keeps its final tour), and
for controlled runs against the per-batch global cost of a full trace. And
the reference's production-event check (test_v2_cuda.py:39-54): timing the
launches changes neither tours nor control states.
"""
import math

import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu

STATIC = {"policy_spec_id": 1, "kind": "static", "mne_level": 0, "max_mne_level": 0, "region": 0,
          "restart_mode": "none", "restart_period": 0, "restart_probability": 0, "stagnation_step": 0,
          "restart_stagnation": 0, "restart_cooldown": 0}


def _cost(xy, tour):
    pts = xy[np.asarray(tour)]
    d = np.sqrt(((pts - np.roll(pts, -1, axis=0)) ** 2).sum(axis=1))
    return int(np.floor(d + 0.5).sum())


def _oracle_settings(oracle_mod, ext, ants):
    s = ext.FixedFacoSettings()
    fields = {k: getattr(s, k) for k in ("primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
                                         "epoch_source_probability", "ls_evaluation_limit",
                                         "initial_ls_evaluation_limit")}
    return oracle_mod.default_settings(ants=ants, **fields)


def test_checkpoints_are_exact_prefixes(ext, oracle_mod):
    xy = uniform_instance(67, seed=607)
    engine = ext.FacoBatchEngine(67, 4, ext.FixedFacoSettings())
    engine.register_problem(1, xy)
    keys, seeds = np.ones(4, dtype=np.uint64), np.array([17, 29, 41, 53], dtype=np.uint64)
    announce("checkpoints-prefix", 67)
    full = engine.evaluate_faco_evaluations(keys, seeds, 64 * 3, checkpoint_iterations=[0, 1, 3])
    assert full["total_tour_evaluations"] == 64 * 3 * 4
    assert [p["iterations"] for p in full["checkpoints"]] == [0, 1, 3]
    previous = [float("inf")] * 4
    settings = _oracle_settings(oracle_mod, ext, 64)
    for point in full["checkpoints"]:
        single = engine.evaluate_faco_evaluations(keys, seeds, 64 * point["iterations"])
        assert point["tours"] == [row["tour"] for row in single["items"]]
        costs = [_cost(xy, tour) for tour in point["tours"]]
        assert all(a <= b for a, b in zip(costs, previous))
        previous = costs
        want = oracle_mod.run_faco(np.repeat(xy[None].astype(np.int32), 4, axis=0), [1] * 4, seeds.tolist(),
                                   settings, point["iterations"], 8)
        assert point["tours"] == [t.tolist() for t in want["tours"]], point["iterations"]
    assert full["checkpoints"][-1]["tours"] == [row["tour"] for row in full["items"]]
    for points in ([1, 1], [3, 1], [4]):
        with pytest.raises(ValueError):
            engine.evaluate_faco_evaluations(keys, seeds, 64 * 3, checkpoint_iterations=points)
    with pytest.raises(ValueError):
        engine.evaluate_faco_evaluations(keys, seeds, 64 * 3, checkpoint_iterations=[-1])


@pytest.mark.parametrize("use_graph", [True, False])
def test_checkpoints_across_graph_segments(ext, oracle_mod, use_graph):
    n, colonies, batches = 300, 6, 40
    xy = clustered_instance(n, seed=3, instance=1)
    engine = ext.FacoBatchEngine(n, colonies, ext.FixedFacoSettings())
    engine.register_problem(9, xy)
    ants = engine.settings.ants
    keys, seeds = np.full(colonies, 9, dtype=np.uint64), np.arange(colonies, dtype=np.uint64) + 100
    points = [5, 16, 17, 33, 40]
    announce("checkpoints-graph", n, use_graph)
    full = engine.evaluate_faco_evaluations(keys, seeds, ants * batches, checkpoint_iterations=points,
                                            use_graph=use_graph)
    plain = engine.evaluate_faco_evaluations(keys, seeds, ants * batches, use_graph=use_graph)
    assert [i["tour"] for i in plain["items"]] == [i["tour"] for i in full["items"]]
    assert full["checkpoints"][-1]["tours"] == [i["tour"] for i in full["items"]]
    settings = _oracle_settings(oracle_mod, ext, ants)
    for point in full["checkpoints"][:2]:
        want = oracle_mod.run_faco(np.repeat(xy[None].astype(np.int32), 2, axis=0), [9, 9], seeds[:2].tolist(),
                                   settings, point["iterations"], 8)
        assert point["tours"][:2] == [t.tolist() for t in want["tours"]], point["iterations"]
    for point in full["checkpoints"][2:]:
        single = engine.evaluate_faco_evaluations(keys, seeds, ants * point["iterations"], use_graph=use_graph)
        assert point["tours"] == [row["tour"] for row in single["items"]], point["iterations"]


def test_mixed_checkpoints_keep_finished_colonies(ext, oracle_mod):
    sizes = (120, 333)   # 64 and 128 ants: 6 and 3 batches
    xys = [uniform_instance(sizes[0], seed=1), clustered_instance(sizes[1], seed=2, instance=1)]
    engine = ext.MixedFacoBatchEngine(max(sizes), 2, ext.FixedFacoSettings())
    for k, xy in enumerate(xys):
        engine.register_problem(k + 1, xy)
    keys, seeds = np.array([1, 2], dtype=np.uint64), np.array([5, 6], dtype=np.uint64)
    limit = math.lcm(64, 128) * 3
    announce("checkpoints-mixed", sizes)
    out = engine.evaluate_faco_evaluations(keys, seeds, limit, checkpoint_iterations=[0, 2, 4, 6])
    for point in out["checkpoints"]:
        for c, (n, ants) in enumerate(zip(sizes, (64, 128))):
            done = min(point["iterations"], limit // ants)
            want = oracle_mod.run_faco(xys[c][None].astype(np.int32), [c + 1], [int(seeds[c])],
                                       _oracle_settings(oracle_mod, ext, ants), done, 8)
            assert len(point["tours"][c]) == n
            assert point["tours"][c] == want["tours"][0].tolist(), (point["iterations"], c)
    with pytest.raises(ValueError):
        engine.evaluate_faco_evaluations(keys, seeds, limit, checkpoint_iterations=[7])


def test_controlled_checkpoints_follow_the_global_cost(ext):
    from tests.test_gpu_control import RESTARTY
    n, panel, batches = 150, 3, 24
    xy = uniform_instance(n, seed=21)
    engine = ext.FacoBatchEngine(n, panel, ext.FixedFacoSettings())
    engine.register_problem(4, xy)
    ants = engine.settings.ants
    keys, seeds = np.full(panel, 4, dtype=np.uint64), np.array([1, 2, 3], dtype=np.uint64)
    points = [0, 1, 7, 16, 24]
    prog = ext.read_program(RESTARTY)
    announce("checkpoints-control", n, batches)
    traced = engine.evaluate_control_raw(keys, seeds, ants * batches, "program", programs=[prog],
                                         trace_batches=list(range(batches)), checkpoint_iterations=points)
    graphed = engine.evaluate_control_raw(keys, seeds, ants * batches, "program", programs=[prog],
                                          checkpoint_iterations=points)
    assert traced["checkpoints"] == graphed["checkpoints"]
    np.testing.assert_array_equal(traced["tours"], graphed["tours"])
    for point in graphed["checkpoints"]:
        k = point["iterations"]
        for c, tour in enumerate(point["tours"]):
            assert sorted(tour) == list(range(n))
            if k:
                assert _cost(xy, tour) == int(traced["trace_costs"][k - 1, c, 1]), (k, c)
    assert graphed["checkpoints"][-1]["tours"] == [t.tolist() for t in graphed["tours"]]
    out = engine.evaluate_program_evaluations(keys, seeds, ants * batches, RESTARTY, checkpoint_iterations=points)
    assert out["checkpoints"] == graphed["checkpoints"]
    base = engine.evaluate_baseline_evaluations(keys, seeds, ants * batches, STATIC,
                                                checkpoint_iterations=[3])
    assert [p["iterations"] for p in base["checkpoints"]] == [3]


@pytest.mark.parametrize("kind", ["program", "baseline"])
def test_production_event_measurement_keeps_tours_and_feedback(ext, kind):
    from tests.test_gpu_control import RESTARTY
    engine = ext.FacoBatchEngine(67, 4, ext.FixedFacoSettings())
    engine.register_problem(1, uniform_instance(67, seed=19))
    keys, seeds = np.ones(4, dtype=np.uint64), np.array([17, 29, 41, 53], dtype=np.uint64)
    controller = RESTARTY if kind == "program" else STATIC
    method = getattr(engine, f"evaluate_{kind}_evaluations")
    announce("events", kind)
    plain = method(keys, seeds, 64 * 3, controller)
    measured = method(keys, seeds, 64 * 3, controller, profile_events_only=True)
    assert [(r["tour"], r["cost"]) for r in plain["items"]] == [(r["tour"], r["cost"]) for r in measured["items"]]
    assert plain["control_states"] == measured["control_states"]
    stages = measured["production_kernel_milliseconds"]
    assert sum(stages.values()) > 0 and stages["construction_and_ls"] > 0 and stages["initialization"] > 0
    faco = engine.evaluate_faco_evaluations(keys, seeds, 64 * 3, profile_events_only=True)
    assert faco["production_kernel_milliseconds"]["construction_and_ls"] > 0
