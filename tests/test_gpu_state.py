"""State snapshots (CountedState) and paired forks of controlled runs on the
device, after the reference's tests/python/test_state_fork_native.py and
tests/cpp/test_state_fork.cpp:

- a capture after k batches plus its continuation equals the uninterrupted
  run bit for bit (tours, costs, control states, every decision, summed work),
  through the serialized bytes and a fresh engine that ran something else in
  between, at cut points 0, 1, 4 and 7 of 7 batches, for programs and
  baselines (periodic restarts keep their batch numbers);
- a one-colony cut of a snapshot continues exactly as that colony did;
- a paired fork forces its first decision, keeps that decision's inputs,
  re-keys the ants, is deterministic, and equals the oracle's fork
  (oracle.run_controlled with fork=...);
- corrupted bytes, foreign settings or coordinates, budgets past the source
  run and invalid fork arguments are rejected.
"""
import json

import numpy as np
import pytest

from gplsaco_amd.instances import uniform_instance
from tests.conftest import announce
from tests.test_counted_state import decode, encode

pytestmark = pytest.mark.gpu

PROGRAM = {"ir_version": 1, "numeric_spec_id": 1, "feature_spec_id": 2,
           "opcode": [0, 0, 2, 0, 0, 3, 0, 4, 2], "operand": [4, 8, 0, 0, 2, 0, 5, 0, 0], "constant_bits": []}


def policy(**kw):
    base = {"policy_spec_id": 1, "kind": "static", "mne_level": 2, "max_mne_level": 2, "region": 1,
            "restart_mode": "none", "restart_period": 0, "restart_probability": 0.0, "stagnation_step": 0,
            "restart_stagnation": 0, "restart_cooldown": 0}
    base.update(kw)
    return base


NAMES = ["xy", "keys", "colony_state", "control_state", "parent", "links", "epoch", "global", "trails", "products",
         "archive_tours", "archive_links", "features", "scores", "masks", "actions", "alternatives", "regions"]


def settings(ext, ants=8):
    s = ext.FixedFacoSettings()
    s.ants = ants
    return s


def maker(ext, n, xys, keys, ants=8):
    def make(colonies):
        e = ext.FacoBatchEngine(n, colonies, settings(ext, ants))
        for k, xy in zip(keys, xys):
            e.register_problem(k, xy)
        return e
    return make


def same_items(a, b):
    assert [i["tour"] for i in a["items"]] == [i["tour"] for i in b["items"]]
    assert [i["cost"] for i in a["items"]] == [i["cost"] for i in b["items"]]
    assert a["control_states"] == b["control_states"]


def same_decisions(rows, want_rows, shift):
    """rows of a continuation (iterations from 1) against the uninterrupted
    run's rows `shift` batches later."""
    assert len(rows) == len(want_rows)
    for row, want in zip(rows, want_rows):
        assert row["iteration"] + shift == want["iteration"] and row["colony"] == want["colony"]
        assert row["action"] == want["action"] and row["legal_mask"] == want["legal_mask"]
        assert np.array_equal(np.float32(row["features"]).view(np.uint32), np.float32(want["features"]).view(np.uint32))
        if "scores" in want:
            assert row["scores"] == want["scores"]


@pytest.mark.parametrize("kind", ["program", "baseline"])
def test_snapshot_rebuilds_the_run(ext, tmp_path, kind):
    n = 67
    xy = uniform_instance(n, seed=763)
    make = maker(ext, n, [xy], [1])
    keys, seeds = np.array([1, 1], dtype=np.uint64), np.array([17, 29], dtype=np.uint64)
    argument = PROGRAM if kind == "program" else policy()
    engine = make(2)
    announce("state-rebuild", kind)
    reference = getattr(engine, f"evaluate_{kind}_evaluations")(keys, seeds, 64, argument)
    captured = getattr(engine, f"capture_{kind}_state")(keys, seeds, 64, 24, argument)
    metadata = captured["state"].describe()
    assert metadata["progress_evaluation_limit"] == 64 and metadata["completed_batches"] == 3
    assert metadata["completed_evaluations_per_colony"] == 24 and metadata["state_spec_id"] == 2
    assert [b["name"] for b in metadata["buffers"]] == NAMES
    assert metadata["tasks"] == [{"instance_key": 1, "seed": 17}, {"instance_key": 1, "seed": 29}]
    assert [m["return_rate"] for m in metadata["members"]] == [
        m["return_rate"] for m in captured["evaluation"]["control_states"]]
    assert [m["incumbent_cost"] for m in metadata["members"]] == [i["cost"] for i in captured["evaluation"]["items"]]
    assert captured["evaluation"]["completed_batches"] == 3
    # the keep regions are the start regions of the next (4th) decision
    raw = engine.evaluate_control_raw(keys, seeds, 64, kind, trace_batches=[3],
                                      **({"programs": [ext.read_program(argument)]} if kind == "program"
                                         else {"baseline": ext.read_baseline(argument)}))
    for c in range(2):
        assert metadata["members"][c]["keep_regions"] == raw["trace_regions"][0, c, 0, :, :16].tolist()
    json.dumps(metadata, allow_nan=False)
    path = tmp_path / "snapshot.bin"
    path.write_bytes(captured["state"].to_bytes())
    serialized = path.read_bytes()
    del engine, captured
    restored = ext.CountedState.from_bytes(path.read_bytes())
    assert restored.describe() == metadata
    engine = make(2)
    engine.evaluate_baseline_evaluations(keys, seeds, 32, policy())   # leaves another run's state behind
    result = getattr(engine, f"continue_{kind}_state")(restored, 40, argument)
    same_items(result, reference)
    assert result["completed_tour_evaluations_per_colony"] == 40 and result["total_tour_evaluations"] == 80
    assert result["budget_seconds"] is None and result["charged_seconds"] == 0
    assert result["overrun_seconds"] == 0 and result["discarded_batches"] == 0
    assert restored.to_bytes() == serialized
    one = getattr(make(1), f"continue_{kind}_state")(restored.select_colony(1), 40, argument)
    assert one["items"][0]["tour"] == reference["items"][1]["tour"]
    assert one["items"][0]["cost"] == reference["items"][1]["cost"]
    assert one["control_states"][0] == reference["control_states"][1]


@pytest.fixture(scope="module")
def four(ext):
    """Four colonies on two instances of 61 cities, 8 ants, 7 batches."""
    n = 61
    xys = [uniform_instance(n, seed=623, instance=i) for i in range(2)]
    keys = [1, 2]
    make = maker(ext, n, xys, keys)
    tasks = (np.array([1, 1, 2, 2], dtype=np.uint64), np.array([17, 29, 17, 29], dtype=np.uint64))
    return make, tasks, xys


def test_cut_points_continue_the_uninterrupted_run(ext, four):
    make, (keys, seeds), _ = four
    ants, batches = 8, 7
    engine = make(4)
    announce("state-cuts")
    every = list(range(1, batches + 1))
    reference = engine.evaluate_program_evaluations(keys, seeds, ants * batches, PROGRAM, decision_iterations=every)
    for cut in (0, 1, 4, 7):
        announce("state-cut", cut)
        prefix = engine.capture_program_state(keys, seeds, ants * batches, ants * cut, PROGRAM,
                                              decision_iterations=every[:cut])
        saved = prefix["state"]
        assert prefix["evaluation"]["completed_batches"] == cut
        assert prefix["evaluation"]["total_tour_evaluations"] == ants * cut * 4
        meta = saved.describe()
        assert meta["completed_batches"] == cut and meta["progress_evaluation_limit"] == ants * batches
        data = saved.to_bytes()
        restored = ext.CountedState.from_bytes(data)
        assert restored.to_bytes() == data and len(meta["buffers"]) == len(NAMES)
        rebuilt = make(4)
        rebuilt.evaluate_baseline_evaluations(keys, seeds, ants * 3, policy())
        suffix = rebuilt.continue_program_state(restored, ants * (batches - cut), PROGRAM,
                                                decision_iterations=list(range(1, batches - cut + 1)))
        for field in ("completed_construction_steps", "completed_ls_evaluations", "total_tour_evaluations"):
            assert prefix["evaluation"][field] + suffix[field] == reference[field], field
        rows = reference["decisions"]
        same_decisions(prefix["evaluation"].get("decisions", []), rows[:4 * cut], 0)
        same_decisions(suffix.get("decisions", []), rows[4 * cut:], cut)
        same_items(suffix, reference)
        assert restored.to_bytes() == data
    # baselines resume their batch numbers: periodic restarts do not count from zero
    periodic = policy(restart_mode="periodic", restart_period=2)
    base_reference = engine.evaluate_baseline_evaluations(keys, seeds, ants * batches, periodic,
                                                          decision_iterations=every)
    assert any(st["restarts"] for st in base_reference["control_states"])
    state = engine.capture_baseline_state(keys, seeds, ants * batches, ants * 3, periodic)["state"]
    suffix = engine.continue_baseline_state(state, ants * 4, periodic, decision_iterations=[1, 2, 3, 4])
    same_items(suffix, base_reference)
    same_decisions(suffix["decisions"], base_reference["decisions"][4 * 3:], 3)


def _outcome(result):
    return [(i["tour"], i["cost"]) for i in result["items"]], result["control_states"], result.get("decisions")


def test_paired_forks(ext, four):
    make, (keys, seeds), _ = four
    ants = 8
    engine = make(4)
    continuation = policy()
    fork_state = engine.capture_program_state(keys, seeds, ants * 7, ants * 4, PROGRAM)["state"]
    original = fork_state.to_bytes()
    solo = make(1)
    selected = fork_state.select_colony(2)
    outcomes = set()
    for region in range(4):
        for seed in (11, 23, 47):
            announce("state-fork", region, seed)
            run = lambda mne, e=engine, s=fork_state: e.continue_baseline_state(  # noqa: E731
                s, ants * 3, continuation, fork_seed=seed, region=region, mne=mne, decision_iterations=[1, 2, 3])
            a, b, again = run(2), run(16), run(2)
            assert _outcome(a) == _outcome(again)
            assert fork_state.to_bytes() == original
            assert a["total_tour_evaluations"] == b["total_tour_evaluations"] == ants * 3 * 4
            first_a, first_b = a["decisions"][:4], b["decisions"][:4]
            assert [r["action"] for r in first_a] == [region * 4] * 4
            assert [r["action"] for r in first_b] == [region * 4 + 3] * 4
            for x, y in zip(first_a, first_b):
                assert x["features"] == y["features"] and x["legal_mask"] == y["legal_mask"]
            assert [r["action"] for r in a["decisions"][4:]] == [r["action"] for r in b["decisions"][4:]]
            outcomes.add(str(_outcome(a)[0]))
            one = solo.continue_baseline_state(selected, ants * 3, continuation, fork_seed=seed, region=region, mne=2,
                                               decision_iterations=[1, 2, 3])
            assert one["items"][0]["tour"] == a["items"][2]["tour"] and one["items"][0]["cost"] == a["items"][2]["cost"]
            assert one["control_states"][0] == a["control_states"][2]
            assert [r["features"] for r in one["decisions"]] == [r["features"] for r in a["decisions"] if r["colony"] == 2]
    assert len(outcomes) > 1       # fork seeds lead to different runs


@pytest.mark.parametrize("region,mne,seed", [(0, 2, 11), (2, 16, 23), (3, 2, 47)])
def test_fork_matches_the_oracle(ext, oracle_mod, region, mne, seed):
    O = oracle_mod
    n, ants, batches, cut = 61, 8, 7, 4
    xys = [uniform_instance(n, seed=624, instance=i) for i in range(2)]
    make = maker(ext, n, xys, [5, 6])
    # colony c runs instance c % 2, as the oracle lays out its panel
    keys, seeds = np.array([5, 6, 5, 6], dtype=np.uint64), np.array([3, 3, 8, 8], dtype=np.uint64)
    base = policy(kind="rule", mne_level=0, max_mne_level=3, region=1, stagnation_step=2, restart_stagnation=5,
                  restart_cooldown=3)
    engine = make(4)
    announce("state-fork-oracle", region, mne, seed)
    prefix = engine.capture_baseline_state(keys, seeds, ants * batches, ants * cut, base)
    suffix = engine.continue_baseline_state(prefix["state"], ants * 3, base, fork_seed=seed, region=region, mne=mne,
                                            decision_iterations=[1, 2, 3])
    s = O.default_settings(ants=ants)
    action = region * 4 + (0 if mne == 2 else 3)
    want = O.run_controlled(np.stack(xys).astype(np.int32), np.array([5, 6], dtype=np.uint64), seeds, s,
                            ants * batches, "baseline", baseline=O.Baseline.from_buffer_copy(bytes(ext.read_baseline(base))),
                            trace=True, stop_batch=cut + 3, fork=(cut, seed, action))
    assert [i["tour"] for i in suffix["items"]] == want["tours"].tolist()
    assert [i["cost"] for i in suffix["items"]] == want["costs"].astype(float).tolist()
    for c in range(4):
        got = suffix["control_states"][c]
        w = want["states"][c]
        assert (got["archive_size"], got["stagnant_batches"], got["epoch_batches"], got["restarts"]) == \
            (w["archive_size"], w["stagnant_batches"], w["epoch_batches"], w["restarts"])
        assert np.float32(got["return_rate"]) == np.float32(w["return_rate"])
        assert np.float32(got["ls_work"]) == np.float32(w["ls_work"])
    steps = prefix["evaluation"]["completed_construction_steps"] + suffix["completed_construction_steps"]
    evals = prefix["evaluation"]["completed_ls_evaluations"] + suffix["completed_ls_evaluations"]
    assert steps == int(want["work"][:, 0].sum()) and evals == int(want["work"][:, 1].sum())
    for row in suffix["decisions"]:
        b, c = cut + row["iteration"] - 1, row["colony"]
        assert row["action"] == int(want["trace_actions"][b, c])
        assert row["legal_mask"] == int(want["trace_masks"][b, c])
        assert np.array_equal(np.float32(row["features"]).view(np.uint32),
                              want["trace_features"][b, :, c, :].view(np.uint32))


def test_invalid_snapshots_and_arguments_are_rejected(ext, four):
    make, (keys, seeds), xys = four
    ants = 8
    engine = make(4)
    continuation = policy()
    announce("state-invalid")
    saved = engine.capture_baseline_state(keys, seeds, ants * 7, ants * 4, continuation)["state"]
    original = saved.to_bytes()
    for value in (b"", original[:-1], original + b"extra"):
        with pytest.raises(ValueError, match="snapshot"):
            ext.CountedState.from_bytes(value)
    for value in (True, -1, 1.5, 1 << 64):
        with pytest.raises(ValueError):
            engine.continue_baseline_state(saved, ants, continuation, fork_seed=value)
        with pytest.raises(ValueError):
            engine.capture_baseline_state(keys, seeds, ants * 7, value, continuation)
    with pytest.raises(ValueError):
        saved.select_colony(True)
    with pytest.raises(ValueError):
        saved.select_colony(4)
    fields = decode(original)
    missing = ext.CountedState.from_bytes(encode(**{**fields, "buffers": fields["buffers"][:-1]}))
    wrong = ext.CountedState.from_bytes(encode(**{**fields, "reals": (fields["reals"][0], 0.6, *fields["reals"][2:])}))
    bad_cost = dict(fields)
    bad_cost["incumbents"] = [(cost + 1, tour) for cost, tour in fields["incumbents"]]
    cost_changed = ext.CountedState.from_bytes(encode(**bad_cost))
    cases = [
        lambda: engine.continue_program_state(missing, ants, PROGRAM),          # a buffer missing
        lambda: engine.continue_program_state(wrong, ants, PROGRAM),            # other settings
        lambda: engine.continue_program_state(cost_changed, ants, PROGRAM),     # host incumbent cost changed
        lambda: engine.continue_program_state(saved, 1, PROGRAM),               # not a whole batch
        lambda: engine.continue_program_state(saved, ants * 4, PROGRAM),        # past the source budget
        lambda: engine.continue_baseline_state(saved, ants, continuation, fork_seed=1, region=0, mne=4),
        lambda: engine.continue_baseline_state(saved, ants, continuation, fork_seed=1, region=4, mne=2),
        lambda: engine.continue_baseline_state(saved, 0, continuation, fork_seed=1),
        lambda: engine.continue_baseline_state(saved, ants, continuation, region=1),   # fork fields without a fork
        lambda: engine.capture_baseline_state(keys, seeds, ants * 7, ants * 8, continuation),  # capture past the budget
        lambda: make(3).continue_program_state(saved, ants, PROGRAM),           # more colonies than the engine
    ]
    for case in cases:
        with pytest.raises(ValueError):
            case()
    other = ext.FacoBatchEngine(61, 4, settings(ext))
    changed = xys[0].copy()
    changed[0, 0] = changed[0, 0] + 1 if changed[0, 0] < 10000 else changed[0, 0] - 1
    other.register_problem(1, changed)
    other.register_problem(2, xys[1])
    with pytest.raises(ValueError, match="coordinates"):
        other.continue_program_state(saved, ants, PROGRAM)
    assert saved.to_bytes() == original
    # the engine still runs after every rejection
    fine = engine.continue_baseline_state(saved, ants * 3, continuation)
    assert fine["completed_batches"] == 3
