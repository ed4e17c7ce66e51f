"""Escape mode on the device (constraint_mode "escape") against the
oracle's Escape runs (oracle.run_faco / run_controlled with escape=, pinned
in tests/test_oracle_escape_mode.py by the reference's replay checks), on
the reference test's graphs (tests/cpp/test_escape_engine.cpp:25-54: the
complete graph on a degenerate instance, and the initial tour plus each
city's four nearest): tours, costs, work, rejections and the eleven escape
counters bit for bit, plain and controlled (a rule baseline that restarts
every batch, so footprints travel through the archive), on the linked-list
(n <= 2048) and the array (n > 2048) construction. With the escape gates
disabled the Escape engine reproduces the Hard engine (the reference's
same(hard, disabled)).
"""
import numpy as np
import pytest

from tests.conftest import announce
from tests.test_oracle_escape_mode import reference_problem

pytestmark = pytest.mark.gpu

FIELDS = ("ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
          "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")


def _setup(ext, O, n, full, ants, colonies, mode="escape"):
    xy, s_oracle, norm, tour, members, rows = reference_problem(n, full, ants=ants)
    s = ext.FixedFacoSettings()
    s.ants = ants
    engine = ext.FacoBatchEngine(n, colonies, s, constraint_mode=mode)
    spec = {"graph_spec_id": 1, "common_initial_tour": tour, "edges": [list(e) for e in sorted(members)],
            "primary": rows[0].tolist(), "backup": rows[1].tolist(), "ls": rows[2].tolist()}
    engine.register_graph_problem(1, xy.astype(np.float64), spec)
    settings = O.default_settings(**{f: getattr(s, f) for f in FIELDS})
    return xy, engine, settings, O.csr(sorted(members), n), rows, members, s


def _cycle(tour):
    return {tuple(sorted((int(tour[i]), int(tour[(i + 1) % len(tour)])))) for i in range(len(tour))}


@pytest.mark.parametrize("n,full,batches,persistent", [(31, True, 16, False), (31, False, 16, False), (200, False, 8, False),
                                                        (1100, False, 3, False), (2100, False, 2, False),
                                                        (31, True, 16, True), (200, False, 8, True), (1100, False, 3, True),
                                                        (2100, False, 2, True)])
def test_escape_runs_match_the_oracle(ext, oracle_mod, n, full, batches, persistent):
    O = oracle_mod
    ants = 4 if n <= 31 else 16
    xy, engine, settings, graph, rows, members, s = _setup(ext, O, n, full, ants, 4)
    keys = np.ones(4, dtype=np.uint64)
    seeds = np.array([17, 29, 31, 37], dtype=np.uint64)
    # one workgroup per colony: the persistent kernel runs every batch; two:
    # per-batch launches of the ant search and the colony update
    engine.set_waves_per_block(engine.launch_shape(4)["waves_per_block"], 1 if persistent else 2)
    assert engine.launch_shape(4)["blocks_per_colony"] == (1 if persistent else 2)
    announce("escape", n, full, persistent)
    got = engine.evaluate_faco_evaluations(keys, seeds, ants * batches)
    assert (got["persistent_launches"] > 0) == persistent, got["persistent_launches"]
    want = O.run_faco(xy[None].astype(np.int32), [1], seeds, settings, batches, 8, graphs=[graph], rows=[rows],
                      escape="on")
    for c in range(4):
        assert got["items"][c]["tour"] == want["tours"][c].tolist(), c
        assert got["items"][c]["cost"] == float(want["costs"][c])
    np.testing.assert_array_equal(got["colony_work"], want["work"])
    np.testing.assert_array_equal(engine.colony_constraint_rejections(4), want["constraint_rejections"])
    np.testing.assert_array_equal(engine.colony_escape_counters(4), want["escape_counters"])
    assert got["constraint_mode"] == "escape" and got["escape_spec_id"] == 1
    assert got["escape_edge_capacity_per_ant"] == 64 and got["reserved_escape_device_bytes"] > 0
    totals = want["escape_counters"].sum(axis=0)
    assert got["escape_counters"] == {k: int(v) for k, v in zip(O.ESCAPE_COUNTERS, totals)}
    if not full:
        assert totals[O.ESCAPE_COUNTERS.index("new_edges")] > 0


@pytest.mark.parametrize("n", [200, 1100])
def test_controlled_escape_runs_match_the_oracle(ext, oracle_mod, n):
    """A static baseline restarting every batch (Bernoulli, probability 1)
    and a GP program that restarts readily: restarts carry archive
    footprints into the parent's, which reactivate cities in the next
    batch's checklists."""
    O = oracle_mod
    ants, batches = 8, 10
    xy, engine, settings, graph, rows, members, s = _setup(ext, O, n, False, ants, 2)
    keys = np.ones(2, dtype=np.uint64)
    seeds = np.array([17, 29], dtype=np.uint64)
    baseline = ext.read_baseline({"policy_spec_id": 1, "kind": "static", "mne_level": 3, "max_mne_level": 3,
                                  "region": 0, "restart_mode": "bernoulli", "restart_period": 0,
                                  "restart_probability": 1.0, "stagnation_step": 0, "restart_stagnation": 0,
                                  "restart_cooldown": 0})
    announce("escape-control", n)
    raw = engine.evaluate_control_raw(keys, seeds, ants * batches, "baseline", baseline=baseline,
                                      trace_batches=range(batches))
    want = O.run_controlled(xy[None].astype(np.int32), np.array([1], dtype=np.uint64), seeds, settings,
                            ants * batches, "baseline", trace=True, graphs=[graph], rows=[rows], escape="on",
                            baseline=O.Baseline.from_buffer_copy(bytes(baseline)))
    assert raw["tours"].tolist() == want["tours"].tolist()
    np.testing.assert_array_equal(raw["work"], want["work"])
    np.testing.assert_array_equal(raw["trace_actions"], want["trace_actions"])
    assert [st["restarts"] for st in raw["states"]] == [st["restarts"] for st in want["states"]]
    np.testing.assert_array_equal(raw["rejections"], want["constraint_rejections"])
    np.testing.assert_array_equal(engine.colony_escape_counters(2), want["escape_counters"])
    old_views = int(want["escape_counters"][:, O.ESCAPE_COUNTERS.index("old_view_reactivations")].sum())
    program = {"ir_version": 1, "numeric_spec_id": 1, "feature_spec_id": 2, "opcode": [0, 0, 2],
               "operand": [4, 1, 0], "constant_bits": []}
    raw = engine.evaluate_control_raw(keys, seeds, ants * batches, "program", programs=[ext.read_program(program)])
    want = O.run_controlled(xy[None].astype(np.int32), np.array([1], dtype=np.uint64), seeds, settings,
                            ants * batches, "program", graphs=[graph], rows=[rows], escape="on",
                            programs=[O.Program.from_buffer_copy(bytes(ext.read_program(program)))])
    assert raw["tours"].tolist() == want["tours"].tolist()
    np.testing.assert_array_equal(raw["work"], want["work"])
    np.testing.assert_array_equal(engine.colony_escape_counters(2), want["escape_counters"])
    old_views += int(want["escape_counters"][:, O.ESCAPE_COUNTERS.index("old_view_reactivations")].sum())
    assert old_views > 0 or n > 1000   # restarts carried archive footprints into a parent


@pytest.mark.parametrize("n,full", [(31, True), (300, False)])
def test_disabled_escape_is_the_hard_run(ext, oracle_mod, n, full):
    """The reference's same(hard, disabled): with every gate shut the Escape
    engine's tours, work and rejections are the Hard engine's, plain and
    controlled, and only the construction opportunities are counted."""
    O = oracle_mod
    ants, batches = 8, 6
    xy, escape, settings, graph, rows, members, s = _setup(ext, O, n, full, ants, 2)
    _, hard, _, _, _, _, _ = _setup(ext, O, n, full, ants, 2, mode="hard")
    keys = np.ones(2, dtype=np.uint64)
    seeds = np.array([5, 6], dtype=np.uint64)
    escape.disable_escape = True
    announce("escape-disabled", n, full)
    a = escape.evaluate_faco_evaluations(keys, seeds, ants * batches)
    b = hard.evaluate_faco_evaluations(keys, seeds, ants * batches)
    assert [it["tour"] for it in a["items"]] == [it["tour"] for it in b["items"]]
    np.testing.assert_array_equal(a["colony_work"], b["colony_work"])
    np.testing.assert_array_equal(escape.colony_constraint_rejections(2), hard.colony_constraint_rejections(2))
    counters = escape.colony_escape_counters(2)
    assert not counters[:, 1:].any() and (counters[:, 0] >= np.asarray(a["colony_work"])[:, 0]).all()
    baseline = ext.read_baseline({"policy_spec_id": 1, "kind": "static", "mne_level": 3, "max_mne_level": 3,
                                  "region": 0, "restart_mode": "bernoulli", "restart_period": 0,
                                  "restart_probability": 1.0, "stagnation_step": 0, "restart_stagnation": 0,
                                  "restart_cooldown": 0})
    ra = escape.evaluate_control_raw(keys, seeds, ants * batches, "baseline", baseline=baseline)
    rb = hard.evaluate_control_raw(keys, seeds, ants * batches, "baseline", baseline=baseline)
    assert ra["tours"].tolist() == rb["tours"].tolist()
    np.testing.assert_array_equal(ra["work"], rb["work"])
    want = O.run_controlled(xy[None].astype(np.int32), np.array([1], dtype=np.uint64), seeds, settings,
                            ants * batches, "baseline", graphs=[graph], rows=[rows], escape="disabled",
                            baseline=O.Baseline.from_buffer_copy(bytes(baseline)))
    assert ra["tours"].tolist() == want["tours"].tolist()


def test_escape_engine_rules(ext, oracle_mod):
    """Escape engines keep the reference's frozen shape (16 / 64 / 20 slots),
    take no time budget, learned prior or continuation, and an unrestricted
    engine can not disable an escape it does not have."""
    O = oracle_mod
    xy, engine, settings, graph, rows, members, s = _setup(ext, O, 200, False, 8, 2)
    wide = ext.FixedFacoSettings()
    wide.ls_width = 24
    with pytest.raises(ValueError):
        ext.FacoBatchEngine(200, 2, wide, constraint_mode="escape")
    keys = np.ones(2, dtype=np.uint64)
    seeds = np.array([1, 2], dtype=np.uint64)
    with pytest.raises(ValueError):
        engine.evaluate(keys, seeds, 0.05, 8)
    with pytest.raises(ValueError):
        engine.set_edge_prior(1, np.ones((200, s.primary_width), dtype=np.float32))
    engine.evaluate_faco_evaluations(keys, seeds, 16)
    with pytest.raises(ValueError):
        engine.evaluate_faco_evaluations(keys, seeds, 16, continue_run=True)
    plain = ext.FacoBatchEngine(200, 2, ext.FixedFacoSettings())
    plain.register_problem(1, xy.astype(np.float64))
    plain.disable_escape = True
    with pytest.raises(ValueError):
        plain.evaluate_faco_evaluations(keys, seeds, 16)
