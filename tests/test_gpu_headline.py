"""Parity of the production shapes against the oracle, bit for bit.

* The headline shape: TSP1000, 128 ants, 256 colonies on one workgroup each,
  which the engine runs as the persistent kernel (faco_colony_run) with 16
  waves per workgroup and the neighbour rows in LDS. Every colony's final
  tour, cost and work counters equal the oracle's.
* Large instances (n in (1500, 2000], 192 ants: the in-place array path of
  construction): fixed-size engines on the persistent and the per-batch
  kernels, mixed-size engines, plain and GP-controlled.
* A slice of the full pipeline: conditional three-tree controllers on a
  mixed TSP1000 + TSP2000 engine with the learned prior's multipliers.

Reference pattern: tests/cpp/test_batch_engine.cpp:39-88 and
tests/cpp/test_control_engine.cpp:70-313 (engine results against the CPU
semantics, colony by colony).
"""
import math

import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from gplsaco_amd.programs import random_controller, random_program
from tests.conftest import announce

pytestmark = pytest.mark.gpu

FIELDS = ("primary_width", "backup_width", "ls_width", "beta", "retention", "p_best", "epoch_source_probability",
          "ls_evaluation_limit", "initial_ls_evaluation_limit")


def _oracle_settings(O, s, ants):
    return O.default_settings(ants=ants, **{k: getattr(s, k) for k in FIELDS})


def _default_ants(n):
    groups = 1
    while 256 * groups * groups < n:
        groups += 1
    return 64 * groups


def _same_as_oracle(got, want, colonies):
    for c in range(colonies):
        assert got["items"][c]["tour"] == want["tours"][c].tolist(), c
        assert got["items"][c]["cost"] == float(want["costs"][c]), c
    np.testing.assert_array_equal(got["colony_work"], want["work"])


@pytest.mark.parametrize("local_search,batches", [("2opt", 20), ("2opt+oropt", 6)])
def test_headline_shape_persistent_kernel_matches_oracle(ext, oracle_mod, local_search, batches):
    n, instances, seeds_per = 1000, 16, 16
    colonies = instances * seeds_per
    s = ext.FixedFacoSettings()
    xys = [(clustered_instance if i % 4 == 3 else uniform_instance)(n, seed=2024, instance=i) for i in range(instances)]
    engine = ext.FacoBatchEngine(n, colonies, s)
    for i, xy in enumerate(xys):
        engine.register_problem(300 + i, xy)
    engine.set_local_search(local_search)
    ants = engine.settings.ants
    assert ants == 128
    shape = engine.launch_shape(colonies)
    # the production shape: one workgroup per colony (the persistent kernel),
    # 16 ants in flight, neighbour rows staged in LDS
    assert (shape["blocks_per_colony"], shape["waves_per_block"], shape["neighbour_rows_in_lds"]) == (1, 16, True), shape
    keys = np.array([300 + c % instances for c in range(colonies)], dtype=np.uint64)
    seeds = np.array([c // instances for c in range(colonies)], dtype=np.uint64)
    announce("headline", local_search, colonies, batches)
    got = engine.evaluate_faco_evaluations(keys, seeds, ants * batches)
    assert got["total_tour_evaluations"] == colonies * ants * batches
    want = oracle_mod.run_faco(np.stack(xys).astype(np.int32), keys[:instances], seeds,
                               _oracle_settings(oracle_mod, s, ants), batches, 8, or_opt=local_search != "2opt")
    _same_as_oracle(got, want, colonies)


def _paths(engine, keys, seeds, limit, split):
    """The persistent kernel (one workgroup per colony) and the per-batch
    kernels (each colony split over `split` workgroups)."""
    waves = engine.launch_shape(keys.size)["waves_per_block"]
    engine.set_waves_per_block(waves, 1)
    assert engine.launch_shape(keys.size)["blocks_per_colony"] == 1
    persistent = engine.evaluate_faco_evaluations(keys, seeds, limit)
    engine.set_waves_per_block(waves, split)
    assert engine.launch_shape(keys.size)["blocks_per_colony"] == split
    per_batch = engine.evaluate_faco_evaluations(keys, seeds, limit)
    engine.set_waves_per_block(0)
    return persistent, per_batch


@pytest.mark.parametrize("n", [1800, 2000])
def test_tsp2000_fixed_engine_both_kernels_match_oracle(ext, oracle_mod, n):
    instances, seeds_per, batches = 2, 3, 4
    colonies = instances * seeds_per
    s = ext.FixedFacoSettings()
    xys = [uniform_instance(n, seed=77, instance=0), clustered_instance(n, seed=77, instance=1)]
    engine = ext.FacoBatchEngine(n, colonies, s)
    for i, xy in enumerate(xys):
        engine.register_problem(900 + i, xy)
    ants = engine.settings.ants
    assert ants == 192
    keys = np.array([900 + c % instances for c in range(colonies)], dtype=np.uint64)
    seeds = np.array([5 + c // instances for c in range(colonies)], dtype=np.uint64)
    announce("tsp-large-fixed", n, colonies, batches)
    persistent, per_batch = _paths(engine, keys, seeds, ants * batches, 2)
    want = oracle_mod.run_faco(np.stack(xys).astype(np.int32), keys[:instances], seeds,
                               _oracle_settings(oracle_mod, s, ants), batches, 8)
    _same_as_oracle(persistent, want, colonies)
    _same_as_oracle(per_batch, want, colonies)


def _mixed_panel(sizes, seed):
    xys = [(clustered_instance if i % 2 else uniform_instance)(n, seed=seed, instance=i) for i, n in enumerate(sizes)]
    return xys, [6000 + i for i in range(len(sizes))]


def test_large_mixed_batch_matches_oracle(ext, oracle_mod):
    sizes = (1800, 2000, 1550, 1024, 1025)
    xys, keys = _mixed_panel(sizes, 31)
    limit = math.lcm(*[_default_ants(n) for n in sizes]) * 2
    colonies = 2 * len(sizes)
    task_keys = np.array([keys[c % len(sizes)] for c in range(colonies)], dtype=np.uint64)
    task_seeds = np.array([1 + c // len(sizes) for c in range(colonies)], dtype=np.uint64)
    engine = ext.MixedFacoBatchEngine(max(sizes), colonies, ext.FixedFacoSettings())
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    announce("tsp-large-mixed", sizes, limit)
    got = engine.evaluate_faco_evaluations(task_keys, task_seeds, limit)
    s = ext.FixedFacoSettings()
    for c in range(colonies):
        t = c % len(sizes)
        ants = _default_ants(sizes[t])
        want = oracle_mod.run_faco(xys[t][None].astype(np.int32), [keys[t]], [int(task_seeds[c])],
                                   _oracle_settings(oracle_mod, s, ants), limit // ants, 8)
        assert got["items"][c]["tour"] == want["tours"][0].tolist(), (c, sizes[t])
        np.testing.assert_array_equal(got["colony_work"][c], want["work"][0])


def _control_against_oracle(ext, oracle_mod, raw, xys, keys, sizes, seeds, limit, mode, deciders,
                            multipliers=None):
    """Colony d * panel + t against the oracle's controlled run of instance t
    alone under decider d, with that instance's own ants and batches."""
    s_default = ext.FixedFacoSettings()
    restarts = 0
    for d, decider in enumerate(deciders):
        for t, n in enumerate(sizes):
            c = d * len(sizes) + t
            ants = _default_ants(n)
            kwargs = {"programs": [oracle_mod.Program.from_buffer_copy(bytes(decider))]} if mode == "program" else \
                {"controllers": [oracle_mod.Controller.from_buffer_copy(bytes(decider))]}
            if multipliers is not None:
                kwargs["multipliers"] = multipliers[t][None]
            want = oracle_mod.run_controlled(xys[t][None].astype(np.int32), [keys[t]], [int(seeds[t])],
                                             _oracle_settings(oracle_mod, s_default, ants), limit, mode, **kwargs)
            assert raw["tours"][c][:n].tolist() == want["tours"][0].tolist(), (d, t)
            assert int(raw["costs"][c]) == int(want["costs"][0]), (d, t)
            np.testing.assert_array_equal(raw["work"][c], want["work"][0])
            for key in ("archive_size", "stagnant_batches", "epoch_batches", "restarts"):
                assert raw["states"][c][key] == want["states"][0][key], (d, t, key)
            assert np.float32(raw["states"][c]["return_rate"]) == np.float32(want["states"][0]["return_rate"])
            assert np.float32(raw["states"][c]["ls_work"]) == np.float32(want["states"][0]["ls_work"])
            restarts += raw["states"][c]["restarts"]
    return restarts


def test_large_mixed_programs_match_oracle(ext, oracle_mod):
    from tests.test_gpu_control import RESTARTY, spec2
    sizes = (1800, 2000)
    xys, keys = _mixed_panel(sizes, 41)
    limit = 192 * 6
    rng = np.random.default_rng(9)
    programs = [ext.read_program(RESTARTY), ext.read_program(spec2(random_program(rng, depth=4)))]
    engine = ext.MixedFacoBatchEngine(max(sizes), len(sizes), ext.FixedFacoSettings(), population_size=2)
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    seeds = np.array([4, 5], dtype=np.uint64)
    announce("tsp-large-programs", sizes, limit)
    raw = engine.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, limit, "program", programs=programs)
    restarts = _control_against_oracle(ext, oracle_mod, raw, xys, keys, sizes, seeds, limit, "program", programs)
    assert restarts > 0, "the restart path was not exercised"


def test_full_pipeline_slice_controllers_with_prior(ext, oracle_mod):
    """Two random conditional three-tree controllers on one TSP1000 and one
    clustered TSP2000 instance of a mixed engine, the committed learned prior
    applied to both (exponent 2), 3 and 2 batches (768 evaluations per
    colony)."""
    from gplsaco_amd import prior as prior_mod
    sizes = (1000, 2000)
    xys = [uniform_instance(1000, seed=515, instance=0), clustered_instance(2000, seed=515, instance=1)]
    keys = [7100, 7101]
    rng = np.random.default_rng(515)
    controllers = [ext.read_controller(random_controller(rng, f"c{i}")) for i in range(2)]
    engine = ext.MixedFacoBatchEngine(2000, len(sizes), ext.FixedFacoSettings(), population_size=2)
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    net = prior_mod.EdgePrior.load()
    announce("c5-slice", sizes)
    multipliers = [engine.apply_edge_prior(net, k) for k in keys]
    seeds = np.array([8, 9], dtype=np.uint64)
    limit = math.lcm(128, 192) * 2
    raw = engine.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, limit, "controller",
                                      controllers=controllers)
    _control_against_oracle(ext, oracle_mod, raw, xys, keys, sizes, seeds, limit, "controller", controllers,
                            multipliers=multipliers)
    # the prior changes the search: plain FACO on the same engine gives other tours
    for k in keys:
        engine.set_edge_prior(k, None)
    plain = engine.evaluate_control_raw(np.array(keys, dtype=np.uint64), seeds, limit, "controller",
                                        controllers=controllers)
    assert plain["tours"].tolist() != raw["tours"].tolist()
