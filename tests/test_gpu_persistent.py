"""The persistent plain-FACO kernel (faco_colony_run: one workgroup per colony
running its ants and its colony update batch after batch) against the
per-batch kernels (an ant launch and a colony-update launch per batch) and
the oracle. Both paths must give the same tours, costs and work counters bit
for bit; the engine takes the persistent path whenever a colony fits one
workgroup (blocks_per_colony 1) and the per-batch path when a colony is split.
"""
import math

import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def _oracle_settings(oracle_mod, ext, ants):
    s = ext.FixedFacoSettings()
    fields = {k: getattr(s, k) for k in ("primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
                                         "epoch_source_probability", "ls_evaluation_limit",
                                         "initial_ls_evaluation_limit")}
    return oracle_mod.default_settings(ants=ants, **fields)


def _both_paths(engine, keys, seeds, limit, ants_split):
    engine.set_waves_per_block(0)
    waves = engine.launch_shape(keys.size)["waves_per_block"]
    engine.set_waves_per_block(waves, 1)            # one workgroup per colony: the persistent kernel
    assert engine.launch_shape(keys.size)["blocks_per_colony"] == 1
    persistent = engine.evaluate_faco_evaluations(keys, seeds, limit)
    engine.set_waves_per_block(waves, ants_split)   # split colonies: the per-batch kernels
    assert engine.launch_shape(keys.size)["blocks_per_colony"] == ants_split
    split = engine.evaluate_faco_evaluations(keys, seeds, limit)
    engine.set_waves_per_block(0)
    return persistent, split


def _same(a, b):
    assert [i["tour"] for i in a["items"]] == [i["tour"] for i in b["items"]]
    assert [i["cost"] for i in a["items"]] == [i["cost"] for i in b["items"]]
    np.testing.assert_array_equal(a["colony_work"], b["colony_work"])


@pytest.mark.parametrize("local_search", ["2opt", "2opt+oropt"])
def test_persistent_equals_per_batch_and_oracle(ext, oracle_mod, local_search):
    n, instances, seeds_per, batches = 160, 4, 64, 9
    colonies = instances * seeds_per
    xys = [(clustered_instance if i % 2 else uniform_instance)(n, seed=5, instance=i) for i in range(instances)]
    engine = ext.FacoBatchEngine(n, colonies, ext.FixedFacoSettings())
    for i, xy in enumerate(xys):
        engine.register_problem(100 + i, xy)
    engine.set_local_search(local_search)
    ants = engine.settings.ants
    keys = np.array([100 + c % instances for c in range(colonies)], dtype=np.uint64)
    seeds = np.array([c // instances for c in range(colonies)], dtype=np.uint64)
    announce("persistent", local_search, colonies, batches)
    persistent, split = _both_paths(engine, keys, seeds, ants * batches, 2)
    _same(persistent, split)
    settings = _oracle_settings(oracle_mod, ext, ants)
    for c in (0, 5, colonies - 1):
        want = oracle_mod.run_faco(xys[c % instances][None].astype(np.int32), [int(keys[c])], [int(seeds[c])],
                                   settings, batches, 8, or_opt=local_search != "2opt")
        assert persistent["items"][c]["tour"] == want["tours"][0].tolist(), c
        np.testing.assert_array_equal(persistent["colony_work"][c], want["work"][0])


def test_persistent_mixed_sizes_stop_at_their_own_budget(ext, oracle_mod):
    sizes = (140, 300, 260)          # 64, 128 and 128 ants per batch
    per_size = 86
    xys = [uniform_instance(sizes[0], seed=8), clustered_instance(sizes[1], seed=8, instance=1),
           uniform_instance(sizes[2], seed=8, instance=2)]
    colonies = per_size * len(sizes)
    engine = ext.MixedFacoBatchEngine(max(sizes), colonies, ext.FixedFacoSettings())
    for i, xy in enumerate(xys):
        engine.register_problem(200 + i, xy)
    keys = np.array([200 + c % len(sizes) for c in range(colonies)], dtype=np.uint64)
    seeds = np.array([c // len(sizes) for c in range(colonies)], dtype=np.uint64)
    limit = math.lcm(64, 128) * 3   # 6 and 3 batches
    announce("persistent-mixed", sizes, colonies)
    persistent, split = _both_paths(engine, keys, seeds, limit, 2)
    _same(persistent, split)
    for c in (0, 1, 2, colonies - 1):
        t = c % len(sizes)
        ants = 64 if sizes[t] <= 256 else 128
        want = oracle_mod.run_faco(xys[t][None].astype(np.int32), [200 + t], [int(seeds[c])],
                                   _oracle_settings(oracle_mod, ext, ants), limit // ants, 8)
        assert persistent["items"][c]["tour"] == want["tours"][0].tolist(), c


@pytest.mark.parametrize("split", [1, 2])
def test_continuation_equals_one_longer_run(ext, split):
    """A run of a batches continued for b more (FACO_RUN_CONTINUE) ends
    exactly where one run of a + b batches ends; its work counters are the
    continuation's own; only the previous plain run of the same tasks can be
    continued."""
    n, instances, seeds_per = 220, 3, 4
    colonies = instances * seeds_per
    xys = [uniform_instance(n, seed=9, instance=i) for i in range(instances)]
    engine = ext.FacoBatchEngine(n, colonies, ext.FixedFacoSettings())
    for i, xy in enumerate(xys):
        engine.register_problem(500 + i, xy)
    ants = engine.settings.ants
    keys = np.array([500 + c % instances for c in range(colonies)], dtype=np.uint64)
    seeds = np.array([c // instances for c in range(colonies)], dtype=np.uint64)
    engine.set_waves_per_block(engine.launch_shape(colonies)["waves_per_block"], split)
    announce("continuation", split)
    whole = engine.evaluate_faco_evaluations(keys, seeds, ants * 7)
    head = engine.evaluate_faco_evaluations(keys, seeds, ants * 3)
    tail = engine.evaluate_faco_evaluations(keys, seeds, ants * 4, continue_run=True)
    assert [i["tour"] for i in tail["items"]] == [i["tour"] for i in whole["items"]]
    assert [i["cost"] for i in tail["items"]] == [i["cost"] for i in whole["items"]]
    np.testing.assert_array_equal(head["colony_work"] + tail["colony_work"], whole["colony_work"])
    assert tail["total_tour_evaluations"] == colonies * ants * 4
    with pytest.raises(ValueError):   # other seeds: not the previous run
        engine.evaluate_faco_evaluations(keys, seeds + 1, ants, continue_run=True)
    engine.evaluate_faco_evaluations(keys, seeds, ants)
    engine.set_local_search("2opt+oropt")   # a changed search forgets the run
    with pytest.raises(ValueError):
        engine.evaluate_faco_evaluations(keys, seeds, ants, continue_run=True)
