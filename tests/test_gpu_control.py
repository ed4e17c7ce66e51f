"""GPU parity of the controlled loop (features, decisions, restarts, archive,
feedback) against the CPU oracle, bit for bit.

The oracle's control layer is pinned to the reference's control_cpu.cpp /
control_ops.hpp / baseline_policy.hpp / factorial_policy.hpp by
tests/test_oracle_control.py. Here every batch's features (FP32 bits),
scores, masks, actions, archive alternatives, start regions and colony costs
are compared with the oracle's, and so are the final tours, costs, work
counters and control summaries.
"""
import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from tests.conftest import announce
from gplsaco_amd.programs import random_program

pytestmark = pytest.mark.gpu

SETTING_FIELDS = ("ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
                  "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")


def settings(ext, **kw):
    s = ext.FixedFacoSettings()
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def oracle_settings(O, s):
    return O.default_settings(**{k: getattr(s, k) for k in SETTING_FIELDS})


def panel(n, count, clustered=False, seed=3):
    gen = clustered_instance if clustered else uniform_instance
    return [gen(n, seed=seed, instance=i) for i in range(count)], [7000 + i for i in range(count)]


def engine_for(ext, n, xys, keys, s, population=1):
    e = ext.FacoBatchEngine(n, len(keys), s, population_size=population)
    for k, xy in zip(keys, xys):
        e.register_problem(k, xy)
    return e


def to_oracle_program(O, p):
    return O.Program.from_buffer_copy(bytes(p))


def to_oracle_controller(O, c):
    return O.Controller.from_buffer_copy(bytes(c))


def to_oracle_baseline(O, b):
    return O.Baseline.from_buffer_copy(bytes(b))


RESTARTY = {"ir_version": 1, "numeric_spec_id": 1, "feature_spec_id": 2, "opcode": [0, 0, 2],
            "operand": [4, 1, 0], "constant_bits": []}   # mode + stagnation: restarts readily


def spec2(p):
    return {**p, "feature_spec_id": 2}


def controller_program(rng, role):
    while True:
        p = spec2(random_program(rng, depth=3))
        bad = any(op == 0 and ((role == 0 and (f == 5 or f >= 8)) or (role == 1 and f == 5))
                  for op, f in zip(p["opcode"], p["operand"]))
        if not bad:
            return p


def compare(raw, want, colonies, batches):
    """Every traced batch and the final state, exactly."""
    for c in range(colonies):
        assert raw["tours"][c].tolist() == want["tours"][c].tolist(), c
    np.testing.assert_array_equal(raw["costs"], want["costs"])
    np.testing.assert_array_equal(raw["work"], want["work"])
    for c in range(colonies):
        got, exp = raw["states"][c], want["states"][c]
        for key in ("archive_size", "stagnant_batches", "epoch_batches", "restarts"):
            assert got[key] == exp[key], (c, key)
        assert np.float32(got["return_rate"]) == np.float32(exp["return_rate"]), c
        assert np.float32(got["ls_work"]) == np.float32(exp["ls_work"]), c
    if "trace_actions" not in raw:
        return
    for b in range(batches):
        f_got, f_want = raw["trace_features"][b], want["trace_features"][b]
        diff = f_got.view(np.uint32) != f_want.view(np.uint32)
        assert not diff.any(), (b, np.argwhere(diff)[:4].tolist())
        np.testing.assert_array_equal(raw["trace_masks"][b], want["trace_masks"][b], err_msg=f"batch {b}")
        np.testing.assert_array_equal(raw["trace_actions"][b], want["trace_actions"][b], err_msg=f"batch {b}")
        np.testing.assert_array_equal(raw["trace_alternatives"][b], want["trace_alternatives"][b])
        np.testing.assert_array_equal(raw["trace_regions"][b], want["trace_regions"][b])
        np.testing.assert_array_equal(raw["trace_costs"][b], want["trace_costs"][b])
        np.testing.assert_array_equal(raw["trace_scores"][b].view(np.uint32), want["trace_scores"][b].view(np.uint32))


@pytest.mark.parametrize("n,count,ants,batches,clustered", [(40, 2, 8, 40, False), (150, 2, 16, 25, True),
                                                            (1000, 1, 64, 6, False)])
def test_program_population_matches_oracle(ext, oracle_mod, n, count, ants, batches, clustered):
    O = oracle_mod
    rng = np.random.default_rng(n)
    xys, keys = panel(n, count, clustered)
    s = settings(ext, ants=ants)
    programs = [RESTARTY, spec2(random_program(rng, depth=4)), RESTARTY]
    e = engine_for(ext, n, xys, keys, s, population=3)
    seeds = np.arange(1, count + 1, dtype=np.uint64)
    parsed = [ext.read_program(p) for p in programs]
    announce("control-program", n, count, ants, batches)
    raw = e.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, ants * batches, "program",
                                 programs=parsed, trace_batches=range(batches))
    C = count * len(programs)
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, np.tile(seeds, len(programs)),
                            oracle_settings(O, s), ants * batches, "program",
                            programs=[to_oracle_program(O, p) for p in parsed], trace=True)
    compare(raw, want, C, batches)
    # identical programs give identical colonies (individual index is no random input)
    assert raw["tours"][:count].tolist() == raw["tours"][2 * count:].tolist()
    assert (raw["trace_actions"] >= 16).any(), "the restart path was not exercised"


def test_conditional_controllers_match_oracle(ext, oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(77)
    n, count, ants, batches = 60, 2, 8, 30
    xys, keys = panel(n, count)
    s = settings(ext, ants=ants)
    dicts = [{"controller_version": 1, "representation": "conditional_three", "controller_id": "c",
              "trees": [controller_program(rng, r) for r in range(3)]} for _ in range(2)]
    dicts.append(RESTARTY)  # a plain program is a one-tree controller
    controllers = [ext.read_controller(d) for d in dicts]
    e = engine_for(ext, n, xys, keys, s, population=3)
    seeds = np.array([3, 4], dtype=np.uint64)
    announce("control-conditional", n)
    raw = e.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, ants * batches, "controller",
                                 controllers=controllers, trace_batches=range(batches))
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, np.tile(seeds, 3), oracle_settings(O, s),
                            ants * batches, "controller", controllers=[to_oracle_controller(O, c) for c in controllers],
                            trace=True)
    compare(raw, want, 6, batches)


BASELINES = [
    dict(kind=0, mne_level=1, max_mne_level=1, region=2, restart_mode=2, restart_probability=0.2),
    dict(kind=0, mne_level=3, max_mne_level=3, region=0, restart_mode=1, restart_period=4),
    dict(kind=1, mne_level=0, max_mne_level=3, region=1, stagnation_step=2, restart_stagnation=5, restart_cooldown=3),
    dict(kind=0, mne_level=2, max_mne_level=2, region=3),
]


def native_baseline(ext, fields):
    from gplsaco_amd._native import BaselinePolicy
    b = BaselinePolicy()
    for k, v in fields.items():
        setattr(b, k, v)
    return b


@pytest.mark.parametrize("policy", range(len(BASELINES)))
def test_baselines_match_oracle(ext, oracle_mod, policy):
    O = oracle_mod
    n, count, ants, batches = 80, 3, 8, 30
    xys, keys = panel(n, count, clustered=True)
    s = settings(ext, ants=ants)
    b = native_baseline(ext, BASELINES[policy])
    e = engine_for(ext, n, xys, keys, s)
    seeds = np.array([11, 12, 13], dtype=np.uint64)
    announce("control-baseline", policy)
    raw = e.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, ants * batches, "baseline", baseline=b,
                                 trace_batches=range(batches))
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, seeds, oracle_settings(O, s), ants * batches,
                            "baseline", baseline=to_oracle_baseline(O, b), trace=True)
    compare(raw, want, count, batches)


@pytest.mark.parametrize("variant", ["M00", "M10", "M01", "M11"])
def test_factorial_variants_match_oracle(ext, oracle_mod, variant):
    O = oracle_mod
    rng = np.random.default_rng(5)
    n, count, ants, batches = 50, 2, 8, 25
    xys, keys = panel(n, count)
    s = settings(ext, ants=ants)
    b = native_baseline(ext, dict(kind=0, mne_level=1, max_mne_level=1, region=1, restart_mode=2,
                                  restart_probability=0.3))
    program = ext.read_program(spec2(random_program(rng, depth=4)))
    e = engine_for(ext, n, xys, keys, s)
    seeds = np.array([21, 22], dtype=np.uint64)
    index = ("M00", "M10", "M01", "M11").index(variant)
    announce("control-factorial", variant)
    raw = e.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, ants * batches, "factorial",
                                 programs=[program], baseline=b, variant=index, trace_batches=range(batches))
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, seeds, oracle_settings(O, s), ants * batches,
                            "factorial", programs=[to_oracle_program(O, program)], baseline=to_oracle_baseline(O, b),
                            variant=variant, trace=True)
    compare(raw, want, count, batches)


def test_forced_fingerprint_collisions_match_oracle(ext, oracle_mod):
    """Every identity reads as zero: duplicates rest on the full comparison
    (the serial archive rule of the update kernel)."""
    O = oracle_mod
    n, count, ants, batches = 30, 2, 8, 20
    xys, keys = panel(n, count)
    s = settings(ext, ants=ants)
    program = ext.read_program(RESTARTY)
    e = engine_for(ext, n, xys, keys, s)
    seeds = np.array([31, 32], dtype=np.uint64)
    announce("control-collisions")
    raw = e.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, ants * batches, "program",
                                 programs=[program], trace_batches=range(batches), force_fingerprint_collisions=True)
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, seeds, oracle_settings(O, s), ants * batches,
                            "program", programs=[to_oracle_program(O, program)], trace=True, force_collisions=True)
    compare(raw, want, count, batches)


def test_graph_replay_equals_eager_and_reference_surface(ext, oracle_mod):
    n, count, ants = 200, 4, 32
    xys, keys = panel(n, count)
    s = settings(ext, ants=ants)
    e = engine_for(ext, n, xys, keys, s, population=2)
    k = np.array(keys, dtype=np.uint64)
    seeds = np.arange(count, dtype=np.uint64)
    rng = np.random.default_rng(9)
    programs = [RESTARTY, spec2(random_program(rng, depth=4))]
    announce("control-graph", n)
    eager = e.evaluate_control_raw(k, seeds, ants * 40, "program", programs=[ext.read_program(p) for p in programs],
                                   use_graph=False)
    graph = e.evaluate_control_raw(k, seeds, ants * 40, "program", programs=[ext.read_program(p) for p in programs],
                                   use_graph=True)
    assert eager["tours"].tolist() == graph["tours"].tolist()
    assert eager["states"] == graph["states"]
    out = e.evaluate_population_evaluations(k, seeds, ants * 40, programs)
    assert len(out) == 2 and out[0]["population_size"] == 2
    assert [i["tour"] for i in out[0]["items"]] == graph["tours"][:count].tolist()
    assert out[1]["control_states"] == graph["states"][count:]
    assert out[0]["total_tour_evaluations"] == ants * 40 * count
    single = ext.FacoBatchEngine(n, count, s)
    for key, xy in zip(keys, xys):
        single.register_problem(key, xy)
    r = single.evaluate_program_evaluations(k, seeds, ants * 40, programs[0], decision_iterations=[1, 7, 40])
    assert [i["tour"] for i in r["items"]] == graph["tours"][:count].tolist()
    assert [row["iteration"] for row in r["decisions"]] == [1] * count + [7] * count + [40] * count
    assert all(len(row["scores"]) == 32 and len(row["features"]) == 12 for row in r["decisions"])
    base = single.evaluate_baseline_evaluations(k, seeds, ants * 16, {
        "policy_spec_id": 1, "kind": "rule", "mne_level": 0, "max_mne_level": 2, "region": 1,
        "restart_mode": "none", "restart_period": 0, "restart_probability": 0, "stagnation_step": 3,
        "restart_stagnation": 6, "restart_cooldown": 2}, decision_iterations=[16])
    assert base["baseline_policy"]["kind"] == "rule" and "scores" not in base["decisions"][0]
    for st in base["control_states"]:
        assert 1 <= st["archive_size"] <= 4


def test_score_controller_gpu_matches_host(ext):
    rng = np.random.default_rng(4)
    features = rng.random((12, 50, 32), dtype=np.float32)
    masks = rng.integers(1, 2**32, size=50, dtype=np.uint32) | 1
    for _ in range(20):
        d = {"controller_version": 1, "representation": "conditional_three", "controller_id": "c",
             "trees": [controller_program(rng, r) for r in range(3)]}
        a = ext.score_controller_cpu(d, features, masks)
        b = ext.score_controller_cuda(d, features, masks)
        assert (a["actions"] == b["actions"]).all()
        assert (a["scores"].view(np.uint32) == b["scores"].view(np.uint32)).all()


@pytest.mark.parametrize("factorial", [None, "M10", "M01"])
def test_population_matches_every_serial_member(ext, factorial):
    """Reference tests/python/test_population_cuda.py:26-80 on integer
    coordinates: every individual of one population launch equals its own
    serial single-program run; chunks of the population and panel changes
    do not change any individual's trajectory."""
    rng = np.random.default_rng(1103)
    n, replicas, iterations = 67, 4, 40
    s = settings(ext)
    values = [spec2(random_program(rng, depth=4)) for _ in range(7)] + [None]
    values[-1] = values[0]
    serial = ext.FacoBatchEngine(n, replicas, s)
    expanded = ext.FacoBatchEngine(n, replicas, s, population_size=len(values))
    for key in (1, 2, 3):
        xy = np.random.default_rng(107 + key).integers(0, 10001, size=(n, 2)).astype(np.float64)
        for engine in (serial, expanded):
            engine.register_problem(key, xy)
    policy = None
    if factorial:
        policy = {"factorial_spec_id": 1, "variant": factorial, "baseline_policy": {
            "policy_spec_id": 1, "kind": "static", "mne_level": 0, "max_mne_level": 0, "region": 0,
            "restart_mode": "none", "restart_period": 0, "restart_probability": 0, "stagnation_step": 0,
            "restart_stagnation": 0, "restart_cooldown": 0}}
    ants = serial.settings.ants
    for panel_keys, panel_seeds in (([1, 1, 2, 2], [17, 29, 17, 29]), ([3, 2, 3, 1], [41, 29, 53, 17])):
        keys = np.asarray(panel_keys, dtype=np.uint64)
        seeds = np.asarray(panel_seeds, dtype=np.uint64)
        announce("population-serial", factorial, panel_keys)
        results = expanded.evaluate_population_evaluations(keys, seeds, iterations * ants, values,
                                                           factorial_policy=policy)
        assert len(results) == len(values)
        for program, actual in zip(values, results):
            expected = (serial.evaluate_factorial_evaluations(keys, seeds, iterations * ants, program, policy)
                        if policy else serial.evaluate_program_evaluations(keys, seeds, iterations * ants, program))
            assert [(r["tour"], r["cost"]) for r in actual["items"]] == [(r["tour"], r["cost"])
                                                                         for r in expected["items"]]
            assert actual["control_states"] == expected["control_states"]
            for field in ("completed_construction_steps", "completed_ls_evaluations",
                          "completed_tour_evaluations_per_colony", "total_tour_evaluations"):
                assert actual[field] == expected[field], field
        assert results[0]["items"] == results[-1]["items"] or \
            [(r["tour"], r["cost"]) for r in results[0]["items"]] == [(r["tour"], r["cost"]) for r in results[-1]["items"]]
        for start, stop in ((3, 8), (0, 1), (0, 3)):
            chunk = expanded.evaluate_population_evaluations(keys, seeds, iterations * ants, values[start:stop],
                                                             factorial_policy=policy)
            for expected, actual in zip(results[start:stop], chunk):
                assert [(r["tour"], r["cost"]) for r in actual["items"]] == [(r["tour"], r["cost"])
                                                                             for r in expected["items"]]
                assert actual["control_states"] == expected["control_states"]


def test_controlled_or_opt_matches_oracle(ext, oracle_mod):
    O = oracle_mod
    n, count, ants, batches = 90, 2, 16, 20
    xys, keys = panel(n, count, clustered=True)
    s = settings(ext, ants=ants)
    e = engine_for(ext, n, xys, keys, s, population=2)
    e.set_local_search("2opt+oropt")
    rng = np.random.default_rng(31)
    parsed = [ext.read_program(RESTARTY), ext.read_program(spec2(random_program(rng, depth=4)))]
    seeds = np.array([2, 3], dtype=np.uint64)
    announce("control-oropt", n)
    raw = e.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, ants * batches, "program",
                                 programs=parsed, trace_batches=range(batches))
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, np.tile(seeds, 2), oracle_settings(O, s),
                            ants * batches, "program", programs=[to_oracle_program(O, p) for p in parsed],
                            trace=True, or_opt=True)
    compare(raw, want, 4, batches)
