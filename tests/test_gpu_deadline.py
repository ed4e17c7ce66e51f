"""Time-budget entry points through the C ABI (faco_engine_evaluate_deadline):
evaluate, evaluate_program and run_program_diagnostic (reference
cpp/src/bindings.cpp:662-700, 937-971; the deadline rules of
cpp/include/gp_faco/deadline.hpp).

Whatever the clock does, the run is a whole number of completed batches:
each incumbent equals the evaluation-count run of that many batches (plain
FACO), or the oracle's controlled run given the progress features the run
decided with (programs). The reference's deadline semantics are checked
directly: nothing is accepted after the budget, a charged preparation counts
against it, a batch that ends late is discarded, and the result dict carries
the reference's keys.
"""
import numpy as np
import pytest

from gplsaco_amd.instances import uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu

KEYS = {"items", "budget_seconds", "elapsed_seconds", "actual_seconds", "charged_seconds",
        "last_batch_completed_seconds", "overrun_seconds", "launched_batches", "completed_batches",
        "discarded_batches", "completed_construction_steps", "completed_ls_evaluations", "allocated_device_bytes",
        "preparation_completed", "preparation_mode"}
FIELDS = ("primary_width", "backup_width", "ls_width", "beta", "retention", "p_best", "epoch_source_probability",
          "ls_evaluation_limit", "initial_ls_evaluation_limit")
SPEC1 = {"ir_version": 1, "numeric_spec_id": 1, "feature_spec_id": 1, "opcode": [0, 0, 2], "operand": [0, 1, 0],
         "constant_bits": []}   # progress + stagnation


def _panel(ext, n=200, instances=2, seeds_per=2, ants=32):
    s = ext.FixedFacoSettings()
    s.ants = ants
    xys = [uniform_instance(n, seed=71, instance=i) for i in range(instances)]
    colonies = instances * seeds_per
    engine = ext.FacoBatchEngine(n, colonies, s)
    for i, xy in enumerate(xys):
        engine.register_problem(800 + i, xy)
    keys = np.array([800 + c % instances for c in range(colonies)], dtype=np.uint64)
    seeds = np.array([1 + c // instances for c in range(colonies)], dtype=np.uint64)
    return engine, xys, keys, seeds, s


def test_plain_deadline_is_whole_batches(ext):
    engine, _, keys, seeds, s = _panel(ext)
    announce("deadline-plain")
    out = engine.evaluate(keys, seeds, 0.25, 8)
    assert KEYS <= set(out) and out["budget_seconds"] == 0.25 and "budget_kind" not in out
    k = out["completed_batches"]
    assert k > 0 and out["preparation_completed"]
    assert out["launched_batches"] == k + out["discarded_batches"] and out["discarded_batches"] <= 1
    assert out["last_batch_completed_seconds"] <= out["elapsed_seconds"]
    for item in out["items"]:
        assert item["has_incumbent"] and item["completed_seconds"] <= 0.25
    # the incumbents are the count run of the same number of batches
    want = engine.evaluate_faco_evaluations(keys, seeds, s.ants * k)
    assert [i["tour"] for i in out["items"]] == [i["tour"] for i in want["items"]]
    assert [i["cost"] for i in out["items"]] == [i["cost"] for i in want["items"]]
    assert out["completed_construction_steps"] == want["completed_construction_steps"]
    assert out["completed_ls_evaluations"] == want["completed_ls_evaluations"]
    # end-to-end preparation: the same search after a fresh preparation
    again = engine.evaluate(keys, seeds, 0.25, 8, "end_to_end")
    assert again["preparation_mode"] == "end_to_end" and again["completed_batches"] > 0
    want = engine.evaluate_faco_evaluations(keys, seeds, s.ants * again["completed_batches"])
    assert [i["tour"] for i in again["items"]] == [i["tour"] for i in want["items"]]


def test_budget_and_charges_bound_what_is_accepted(ext):
    engine, _, keys, seeds, _ = _panel(ext)
    announce("deadline-charges")
    none = engine.evaluate(keys, seeds, 0.0, 8)
    assert not none["preparation_completed"] and none["completed_batches"] == 0
    assert all(not i["has_incumbent"] and i["cost"] is None for i in none["items"])
    # charges: the cheap tour fits, the preparation does not
    for k in (800, 801):
        engine.set_preparation_charges(k, 0.01, 30.0)
    with pytest.raises(ValueError):
        engine.set_preparation_charges(800, 0.02, 30.0)   # frozen
    out = engine.evaluate(keys, seeds, 1.0, 8)
    assert not out["preparation_completed"] and out["completed_batches"] == 0
    assert out["charged_seconds"] >= 30.0 and out["overrun_seconds"] > 0
    for c, item in enumerate(out["items"]):
        data = engine.problem_data(int(keys[c]))
        assert item["has_incumbent"] and len(item["tour"]) == 200
        assert item["cost"] > data["initial_cost"]   # the cheap tour, not the prepared one


def _oracle_settings(O, s):
    return O.default_settings(**{k: getattr(s, k) for k in ("ants",) + FIELDS})


def _against_oracle(ext, O, out, xys, keys, seeds, s, program):
    k = out["completed_batches"]
    assert k >= 1
    want = O.run_controlled(np.stack(xys).astype(np.int32), [int(x) for x in keys[:len(xys)]], seeds,
                            _oracle_settings(O, s), s.ants * k, "program",
                            programs=[O.Program.from_buffer_copy(bytes(ext.read_program(program)))],
                            progress=out["progress_features"][:k])
    for c, item in enumerate(out["items"]):
        assert item["tour"] == want["tours"][c].tolist(), c
        assert item["cost"] == float(want["costs"][c]), c
    for c, st in enumerate(out["control_states"]):
        for key in ("archive_size", "stagnant_batches", "epoch_batches", "restarts"):
            assert st[key] == want["states"][c][key], (c, key)


def test_program_under_a_deadline_matches_the_oracle(ext, oracle_mod):
    engine, xys, keys, seeds, s = _panel(ext)
    announce("deadline-program")
    out = engine.evaluate_program(keys, seeds, 0.3, SPEC1)
    k = out["completed_batches"]
    assert k > 1 and len(out["progress_features"]) == out["launched_batches"]
    ratios = out["progress_features"]
    assert all(0 < a < b for a, b in zip(ratios, ratios[1:]))    # the clock advances
    _against_oracle(ext, oracle_mod, out, xys, keys, seeds, s, SPEC1)
    with pytest.raises(ValueError):   # evaluation-count programs (feature spec 2) are not time programs
        engine.evaluate_program(keys, seeds, 0.3, {**SPEC1, "feature_spec_id": 2})


def test_program_diagnostic_fixed_batches(ext, oracle_mod):
    engine, xys, keys, seeds, s = _panel(ext)
    announce("deadline-diagnostic")
    out = engine.run_program_diagnostic(keys, seeds, SPEC1, 6, elapsed_ratio=0.25, profile=True)
    assert out["scope"] == "fixed_batch_diagnostic_not_fitness"
    assert out["completed_batches"] == 6 and out["progress_features"] == [0.25] * 6
    rows = out["profile"]["batches"]
    assert [r["batch"] for r in rows] == list(range(6)) and all(r["committed"] for r in rows)
    assert all(r["gpu_milliseconds"]["construction_and_ls"] > 0 for r in rows)
    _against_oracle(ext, oracle_mod, out, xys, keys, seeds, s, SPEC1)


def test_deadline_argument_types_are_strict(ext):
    engine, _, keys, seeds, _ = _panel(ext)
    with pytest.raises(ValueError):
        engine.evaluate(keys, seeds, True, 8)
    with pytest.raises(ValueError):
        engine.evaluate(keys, seeds, 1.0, 8.0)
    with pytest.raises(ValueError):
        engine.evaluate(keys, seeds, 1.0, 8, "cached")
    with pytest.raises(ValueError):
        engine.evaluate(keys, seeds, float("nan"), 8)
    with pytest.raises(TypeError):   # arrays are taken without conversion
        engine.evaluate(keys.astype(np.int64), seeds, 1.0, 8)
    with pytest.raises(ValueError):
        engine.run_program_diagnostic(keys, seeds, SPEC1, 0)
    with pytest.raises(ValueError):
        engine.run_program_diagnostic(keys, seeds, SPEC1, 3, elapsed_ratio=1.5)
