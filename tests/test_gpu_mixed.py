"""Mixed-size batches: instances of different n in one engine and one launch
per iteration. Each colony must equal the oracle's FACO run of its instance
alone, with its own ants (the per-size default) and its own number of
iterations, bit for bit (tour, cost, work counters). The autotuner may change
the launch shape but never a result.
"""
import math

import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def _oracle_settings(oracle_mod, s, ants):
    fields = {k: getattr(s, k) for k in ("primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
                                         "epoch_source_probability", "ls_evaluation_limit",
                                         "initial_ls_evaluation_limit")}
    return oracle_mod.default_settings(ants=ants, **fields)


def _default_ants(n):
    groups = 1
    while 256 * groups * groups < n:
        groups += 1
    return 64 * groups


def _panel(sizes):
    xys = [(clustered_instance if i % 2 else uniform_instance)(n, seed=11, instance=i) for i, n in enumerate(sizes)]
    keys = [5000 + i for i in range(len(sizes))]
    return xys, keys


def _check_against_oracle(oracle_mod, ext, got, xys, keys, task_keys, task_seeds, limit, or_opt=False):
    s = ext.FixedFacoSettings()
    for c, (k, seed) in enumerate(zip(task_keys, task_seeds)):
        i = keys.index(int(k))
        n = xys[i].shape[0]
        ants = _default_ants(n)
        want = oracle_mod.run_faco(xys[i][None].astype(np.int32), [int(k)], [int(seed)],
                                   _oracle_settings(oracle_mod, s, ants), limit // ants, 8, or_opt=or_opt)
        assert len(got["items"][c]["tour"]) == n
        assert got["items"][c]["tour"] == want["tours"][0].tolist(), (c, n)
        assert got["items"][c]["cost"] == float(want["costs"][0]), (c, n)
        np.testing.assert_array_equal(got["colony_work"][c], want["work"][0])


@pytest.mark.parametrize("sizes,seeds,rounds", [
    ((120, 200, 333, 500), 2, 3),
    ((1300, 257, 1000, 180, 700), 1, 2),
])
def test_mixed_batch_matches_oracle_per_colony(ext, oracle_mod, sizes, seeds, rounds):
    xys, keys = _panel(sizes)
    lcm = math.lcm(*[_default_ants(n) for n in sizes])
    limit = lcm * rounds
    colonies = len(sizes) * seeds
    task_keys = np.array([keys[c % len(sizes)] for c in range(colonies)], dtype=np.uint64)
    task_seeds = np.array([7 + c // len(sizes) for c in range(colonies)], dtype=np.uint64)
    engine = ext.MixedFacoBatchEngine(max(sizes), colonies, ext.FixedFacoSettings())
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    announce("mixed", sizes, colonies, limit)
    got = engine.evaluate_faco_evaluations(task_keys, task_seeds, limit, use_graph=True)
    assert got["completed_tour_evaluations_per_colony"] == limit
    assert got["total_tour_evaluations"] == limit * colonies
    _check_against_oracle(oracle_mod, ext, got, xys, keys, task_keys, task_seeds, limit)
    # eager launches give the same results as the graph replay
    again = engine.evaluate_faco_evaluations(task_keys, task_seeds, limit, use_graph=False)
    assert [i["tour"] for i in again["items"]] == [i["tour"] for i in got["items"]]


def test_mixed_batch_or_opt_and_prior(ext, oracle_mod):
    sizes = (150, 400, 260)
    xys, keys = _panel(sizes)
    limit = math.lcm(*[_default_ants(n) for n in sizes]) * 2
    engine = ext.MixedFacoBatchEngine(max(sizes), len(sizes), ext.FixedFacoSettings())
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    engine.set_local_search("2opt+oropt")
    task_keys = np.array(keys, dtype=np.uint64)
    task_seeds = np.array([3, 4, 5], dtype=np.uint64)
    announce("mixed-oropt", sizes)
    got = engine.evaluate_faco_evaluations(task_keys, task_seeds, limit, use_graph=False)
    _check_against_oracle(oracle_mod, ext, got, xys, keys, task_keys, task_seeds, limit, or_opt=True)
    # per-instance prior multipliers at each instance's own size
    engine.set_local_search("2opt")
    rng = np.random.default_rng(1)
    mults = {}
    for k, n in zip(keys, sizes):
        mults[k] = rng.uniform(0.05, 1.0, size=(n, engine.settings.primary_width)).astype(np.float32)
        engine.set_edge_prior(k, mults[k])
        assert engine.edge_features(k).shape == (n * engine.settings.primary_width, 8)
    got = engine.evaluate_faco_evaluations(task_keys, task_seeds, limit, use_graph=False)
    s = ext.FixedFacoSettings()
    for c, k in enumerate(keys):
        ants = _default_ants(sizes[c])
        want = oracle_mod.run_faco(xys[c][None].astype(np.int32), [k], [int(task_seeds[c])],
                                   _oracle_settings(oracle_mod, s, ants), limit // ants, 8,
                                   multipliers=mults[k][None])
        assert got["items"][c]["tour"] == want["tours"][0].tolist(), c


def test_autotune_keeps_results_and_picks_the_fastest(ext):
    sizes = (300, 900, 1500, 220)
    xys, keys = _panel(sizes)
    limit = math.lcm(*[_default_ants(n) for n in sizes]) * 2
    colonies = 8
    task_keys = np.array([keys[c % 4] for c in range(colonies)], dtype=np.uint64)
    task_seeds = np.array([c // 4 for c in range(colonies)], dtype=np.uint64)
    engine = ext.MixedFacoBatchEngine(max(sizes), colonies, ext.FixedFacoSettings())
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    announce("autotune", sizes)
    before = engine.evaluate_faco_evaluations(task_keys, task_seeds, limit)
    tuned = engine.autotune(task_keys, task_seeds, iterations=2)
    rows = tuned["table"]
    assert len(rows) >= 2
    assert tuned["best"]["ms_per_iteration"] == min(r["ms_per_iteration"] for r in rows)
    assert tuned["launch_shape"]["waves_per_block"] == tuned["best"]["waves"]
    after = engine.evaluate_faco_evaluations(task_keys, task_seeds, limit)
    assert [i["tour"] for i in after["items"]] == [i["tour"] for i in before["items"]]
    assert [i["cost"] for i in after["items"]] == [i["cost"] for i in before["items"]]


def test_mixed_engine_rejects_bad_inputs(ext):
    engine = ext.MixedFacoBatchEngine(500, 2, ext.FixedFacoSettings())
    with pytest.raises(ValueError):
        engine.register_problem(1, uniform_instance(600, seed=1))
    with pytest.raises(ValueError):
        engine.register_problem(2, uniform_instance(40, seed=1))   # narrower candidate rows
    engine.register_problem(3, uniform_instance(200, seed=1))
    engine.register_problem(4, uniform_instance(500, seed=1))
    with pytest.raises(ValueError):   # 64 and 128 ants: 64 * 3 is not a whole number of 128-ant batches
        engine.evaluate_faco_evaluations(np.array([3, 4], np.uint64), np.array([0, 0], np.uint64), 192)
    from tests.test_gpu_control import RESTARTY
    with pytest.raises(ValueError):   # the same whole-batch rule for controlled runs
        engine.evaluate_program_evaluations(np.array([3, 4], np.uint64), np.array([0, 0], np.uint64), 192, RESTARTY)


def test_mixed_program_population_matches_oracle(ext, oracle_mod):
    """GP-controlled FACO on a mixed-size panel (the uneven-batch shape at a
    small scale): colony d * panel + t equals the oracle's controlled run of
    instance t alone under program d, with that instance's own ants and
    iteration count (features, actions, restarts, archive, tours)."""
    from gplsaco_amd.programs import random_program
    from tests.test_gpu_control import RESTARTY, spec2, to_oracle_program
    sizes = (130, 300, 220)
    xys, keys = _panel(sizes)
    limit = math.lcm(*[_default_ants(n) for n in sizes]) * 8
    rng = np.random.default_rng(5)
    programs = [RESTARTY, spec2(random_program(rng, depth=4))]
    engine = ext.MixedFacoBatchEngine(max(sizes), len(sizes), ext.FixedFacoSettings(), population_size=2)
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    parsed = [ext.read_program(p) for p in programs]
    seeds = np.array([2, 3, 4], dtype=np.uint64)
    announce("mixed-control", sizes, limit)
    raw = engine.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, limit, "program", programs=parsed)
    s = ext.FixedFacoSettings()
    restarts = 0
    for d, prog in enumerate(parsed):
        for t, n in enumerate(sizes):
            c = d * len(sizes) + t
            ants = _default_ants(n)
            want = oracle_mod.run_controlled(xys[t][None].astype(np.int32), [keys[t]], [int(seeds[t])],
                                             _oracle_settings(oracle_mod, s, ants), limit, "program",
                                             programs=[to_oracle_program(oracle_mod, prog)])
            assert raw["tours"][c][:n].tolist() == want["tours"][0].tolist(), (d, t)
            assert (raw["tours"][c][n:] == 0xFFFFFFFF).all()
            assert int(raw["costs"][c]) == int(want["costs"][0]), (d, t)
            np.testing.assert_array_equal(raw["work"][c], want["work"][0])
            for key in ("archive_size", "stagnant_batches", "epoch_batches", "restarts"):
                assert raw["states"][c][key] == want["states"][0][key], (d, t, key)
            assert np.float32(raw["states"][c]["return_rate"]) == np.float32(want["states"][0]["return_rate"])
            restarts += raw["states"][c]["restarts"]
    assert restarts > 0, "the restart path was not exercised"
    out = engine.evaluate_program_evaluations(np.array(keys, dtype=np.uint64), seeds, limit, programs[1])
    assert [len(item["tour"]) for item in out["items"]] == list(sizes)
