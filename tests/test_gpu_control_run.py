"""The persistent controlled kernel (faco_control_run): one launch runs every
batch of a GP-controlled population, each colony's workgroup doing the
decision, its ants, the colony update and the archive / feedback update in
turn (the per-batch kernels' order, reference cuda/batch_engine.cu:880-1060).

It must equal, bit for bit, the per-batch launches of the same engine
(tours, costs, work counters, every control summary) and the oracle's
controlled run, for programs, conditional controllers, rule baselines,
factorial masks, Or-opt, neighbour rows in LDS or in global memory, the
linked-list (n <= 2048) and the array construction, mixed-size panels and
runs split by checkpoints. `persistent_launches` shows which path ran.
"""
import math

import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from gplsaco_amd.programs import random_program
from tests.conftest import announce
from tests.test_gpu_control import (RESTARTY, controller_program, oracle_settings, panel, settings, spec2,
                                    to_oracle_baseline, to_oracle_controller, to_oracle_program)

pytestmark = pytest.mark.gpu


def _same(a, b, colonies):
    assert a["tours"].tolist() == b["tours"].tolist()
    np.testing.assert_array_equal(a["costs"], b["costs"])
    np.testing.assert_array_equal(a["work"], b["work"])
    assert a["states"] == b["states"]
    assert len(a["states"]) == colonies


def _against_oracle(raw, want, colonies):
    assert raw["tours"].tolist() == want["tours"].tolist()
    np.testing.assert_array_equal(raw["costs"], want["costs"])
    np.testing.assert_array_equal(raw["work"], want["work"])
    for c in range(colonies):
        for key in ("archive_size", "stagnant_batches", "epoch_batches", "restarts"):
            assert raw["states"][c][key] == want["states"][c][key], (c, key)
        assert np.float32(raw["states"][c]["return_rate"]) == np.float32(want["states"][c]["return_rate"]), c
        assert np.float32(raw["states"][c]["ls_work"]) == np.float32(want["states"][c]["ls_work"]), c


def _engine(ext, n, count, ants, population, waves, clustered=False):
    xys, keys = panel(n, count, clustered)
    s = settings(ext, ants=ants)
    e = ext.FacoBatchEngine(n, count, s, population_size=population)
    for k, xy in zip(keys, xys):
        e.register_problem(k, xy)
    e.set_waves_per_block(waves, 1)      # one workgroup per colony: the persistent path
    return e, xys, keys, s


def _both(e, keys, seeds, limit, mode, **kw):
    k = np.array(keys, dtype=np.uint64)
    fused = e.evaluate_control_raw(k, seeds, limit, mode, use_graph=True, **kw)
    eager = e.evaluate_control_raw(k, seeds, limit, mode, use_graph=False, **kw)
    assert fused["result"].persistent_launches >= 1, "the persistent controlled kernel did not run"
    assert eager["result"].persistent_launches == 0
    return fused, eager


@pytest.mark.parametrize("n,ants,batches,waves", [(200, 32, 30, 8), (90, 16, 40, 2), (1100, 64, 5, 12),
                                                  (300, 32, 12, 6 | 0x100), (400, 64, 10, 16), (2100, 32, 3, 4 | 0x100)])
def test_program_population(ext, oracle_mod, n, ants, batches, waves):
    O = oracle_mod
    count = 2
    e, xys, keys, s = _engine(ext, n, count, ants, 3, waves, clustered=n % 2 == 1)
    rng = np.random.default_rng(n)
    programs = [ext.read_program(p) for p in (RESTARTY, spec2(random_program(rng, depth=4)), RESTARTY)]
    seeds = np.arange(1, count + 1, dtype=np.uint64)
    announce("control-run-program", n, ants, batches, waves)
    fused, eager = _both(e, keys, seeds, ants * batches, "program", programs=programs)
    _same(fused, eager, 3 * count)
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, np.tile(seeds, 3), oracle_settings(O, s),
                            ants * batches, "program", programs=[to_oracle_program(O, p) for p in programs])
    _against_oracle(fused, want, 3 * count)
    assert sum(st["restarts"] for st in fused["states"]) > 0, "the restart path was not exercised"


def test_conditional_controllers(ext, oracle_mod):
    O = oracle_mod
    n, count, ants, batches = 150, 2, 16, 30
    e, xys, keys, s = _engine(ext, n, count, ants, 3, 4)
    rng = np.random.default_rng(78)
    dicts = [{"controller_version": 1, "representation": "conditional_three", "controller_id": "c",
              "trees": [controller_program(rng, r) for r in range(3)]} for _ in range(2)] + [RESTARTY]
    controllers = [ext.read_controller(d) for d in dicts]
    seeds = np.array([3, 4], dtype=np.uint64)
    announce("control-run-controllers", n)
    fused, eager = _both(e, keys, seeds, ants * batches, "controller", controllers=controllers)
    _same(fused, eager, 6)
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, np.tile(seeds, 3), oracle_settings(O, s),
                            ants * batches, "controller", controllers=[to_oracle_controller(O, c) for c in controllers])
    _against_oracle(fused, want, 6)


BASELINE = dict(kind=1, mne_level=0, max_mne_level=3, region=1, stagnation_step=2, restart_stagnation=5,
                restart_cooldown=3)


@pytest.mark.parametrize("mode,variant", [("baseline", 3), ("factorial", 1), ("factorial", 2)])
def test_baseline_and_factorial(ext, oracle_mod, mode, variant):
    from gplsaco_amd._native import BaselinePolicy
    O = oracle_mod
    n, count, ants, batches = 120, 2, 16, 36
    population = 2 if mode == "factorial" else 1
    e, xys, keys, s = _engine(ext, n, count, ants, population, 4)
    b = BaselinePolicy()
    for k, v in BASELINE.items():
        setattr(b, k, v)
    rng = np.random.default_rng(variant)
    kw = {"baseline": b}
    okw = {"baseline": to_oracle_baseline(O, b)}
    if mode == "factorial":
        progs = [ext.read_program(p) for p in (RESTARTY, spec2(random_program(rng, depth=4)))]
        kw.update(programs=progs, variant=variant)
        okw.update(programs=[to_oracle_program(O, p) for p in progs], variant={1: "M10", 2: "M01"}[variant])
    seeds = np.array([5, 6], dtype=np.uint64)
    announce("control-run", mode, variant)
    fused, eager = _both(e, keys, seeds, ants * batches, mode, **kw)
    _same(fused, eager, population * count)
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, np.tile(seeds, population), oracle_settings(O, s),
                            ants * batches, mode, **okw)
    _against_oracle(fused, want, population * count)


def test_or_opt_and_checkpoints(ext, oracle_mod):
    """Or-opt ants, and a run cut into three persistent launches by its
    checkpoints: the same result and the same checkpoint tours."""
    O = oracle_mod
    n, count, ants, batches = 110, 2, 16, 24
    e, xys, keys, s = _engine(ext, n, count, ants, 2, 8, clustered=True)
    e.set_local_search("2opt+oropt")
    rng = np.random.default_rng(32)
    programs = [ext.read_program(p) for p in (RESTARTY, spec2(random_program(rng, depth=4)))]
    seeds = np.array([2, 3], dtype=np.uint64)
    announce("control-run-oropt")
    fused, eager = _both(e, keys, seeds, ants * batches, "program", programs=programs,
                         checkpoint_iterations=(5, 17))
    assert fused["result"].persistent_launches == 3
    _same(fused, eager, 4)
    assert [c["iterations"] for c in fused["checkpoints"]] == [5, 17]
    assert fused["checkpoints"] == eager["checkpoints"]
    want = O.run_controlled(np.stack(xys).astype(np.int32), keys, np.tile(seeds, 2), oracle_settings(O, s),
                            ants * batches, "program", programs=[to_oracle_program(O, p) for p in programs],
                            or_opt=True)
    _against_oracle(fused, want, 4)


def _default_ants(n):
    groups = 1
    while 256 * groups * groups < n:
        groups += 1
    return 64 * groups


def test_mixed_sizes(ext, oracle_mod):
    """A mixed-size population in one persistent launch: every colony stops
    at its own budget and equals the oracle's run of its instance alone."""
    from tests.test_gpu_mixed import _oracle_settings
    sizes = (130, 420, 260, 1030)
    xys = [(clustered_instance if i % 2 else uniform_instance)(n, seed=13, instance=i) for i, n in enumerate(sizes)]
    keys = [6100 + i for i in range(len(sizes))]
    limit = math.lcm(*[_default_ants(n) for n in sizes]) * 6
    rng = np.random.default_rng(6)
    programs = [ext.read_program(p) for p in (RESTARTY, spec2(random_program(rng, depth=4)))]
    e = ext.MixedFacoBatchEngine(max(sizes), len(sizes), ext.FixedFacoSettings(), population_size=2)
    for k, xy in zip(keys, xys):
        e.register_problem(k, xy)
    e.set_waves_per_block(10, 1)
    seeds = np.array([2, 3, 4, 5], dtype=np.uint64)
    announce("control-run-mixed", sizes, limit)
    fused, eager = _both(e, keys, seeds, limit, "program", programs=programs)
    _same(fused, eager, 2 * len(sizes))
    s = ext.FixedFacoSettings()
    for d, prog in enumerate(programs):
        for t, n in enumerate(sizes):
            c = d * len(sizes) + t
            want = oracle_mod.run_controlled(xys[t][None].astype(np.int32), [keys[t]], [int(seeds[t])],
                                             _oracle_settings(oracle_mod, s, _default_ants(n)), limit, "program",
                                             programs=[to_oracle_program(oracle_mod, prog)])
            assert fused["tours"][c][:n].tolist() == want["tours"][0].tolist(), (d, t)
            assert int(fused["costs"][c]) == int(want["costs"][0]), (d, t)
            np.testing.assert_array_equal(fused["work"][c], want["work"][0])
            for key in ("archive_size", "stagnant_batches", "epoch_batches", "restarts"):
                assert fused["states"][c][key] == want["states"][0][key], (d, t, key)


def test_population_entry_point_uses_it(ext):
    """evaluate_population_evaluations on a one-workgroup-per-colony shape
    runs the persistent kernel and still equals each member's serial run."""
    n, count, ants = 140, 2, 32
    e, xys, keys, s = _engine(ext, n, count, ants, 3, 8)
    serial, _, _, _ = _engine(ext, n, count, ants, 1, 8)
    rng = np.random.default_rng(8)
    programs = [RESTARTY, spec2(random_program(rng, depth=4)), spec2(random_program(rng, depth=3))]
    k = np.array(keys, dtype=np.uint64)
    seeds = np.array([7, 8], dtype=np.uint64)
    announce("control-run-population")
    out = e.evaluate_population_evaluations(k, seeds, ants * 20, programs)
    assert all(o["persistent_launches"] == 1 for o in out)
    for program, got in zip(programs, out):
        want = serial.evaluate_program_evaluations(k, seeds, ants * 20, program)
        assert want["persistent_launches"] == 1
        assert [(i["tour"], i["cost"]) for i in got["items"]] == [(i["tour"], i["cost"]) for i in want["items"]]
        assert got["control_states"] == want["control_states"]


def test_controlled_autotune_keeps_results(ext):
    """The controlled autotuner on a fresh engine times every shape both as
    the persistent kernel and as per-batch launches, keeps the fastest, and
    the run after it equals the same run on the default shape."""
    n, count, ants = 160, 2, 32
    xys, keys = panel(n, count)
    s = settings(ext, ants=ants)
    rng = np.random.default_rng(12)
    programs = [ext.read_program(p) for p in (RESTARTY, spec2(random_program(rng, depth=4)))]
    k = np.array(keys, dtype=np.uint64)
    seeds = np.array([1, 2], dtype=np.uint64)
    tuned = ext.FacoBatchEngine(n, count, s, population_size=2)
    plain = ext.FacoBatchEngine(n, count, s, population_size=2)
    for key, xy in zip(keys, xys):
        tuned.register_problem(key, xy)
        plain.register_problem(key, xy)
    announce("control-run-autotune")
    out = tuned.autotune_control(k, seeds, "program", programs=programs, iterations=2)
    assert any(r["persistent"] for r in out["table"]) and any(not r["persistent"] for r in out["table"])
    assert out["best"] == min(out["table"], key=lambda r: r["ms_per_iteration"])
    got = tuned.evaluate_control_raw(k, seeds, ants * 12, "program", programs=programs)
    assert (got["result"].persistent_launches > 0) == out["best"]["persistent"]
    want = plain.evaluate_control_raw(k, seeds, ants * 12, "program", programs=programs, use_graph=False)
    _same(got, want, 4)
