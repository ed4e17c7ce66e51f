"""Headline benchmark: ant tours/s of plain FACO at TSP1000 on one MI355X.

Workload: 32 uniform random Euclidean TSP1000 instances (integer coordinates
in [0,10000]^2, EUC_2D) x 8 seeds = 256 colonies, 128 ants per colony, MNE 8,
16/64/20 candidate widths, 2-opt checklist local search with the 100000
evaluation limit: a 100-iteration run from the prepared initial tour (the
length of the reference's TSP1000 profile row). A step is one FACO iteration
of every colony: construction + local search for all 32768 ants, then
selection and the pheromone update. The timed steps are the LAST --steps
iterations of that run (the steady state: with --steps 20, iterations
81-100); the iterations before them are the warm-up (at least --warmup),
run through the per-batch kernels, which give bit-identical colony states,
so that the production kernel's only dispatch is the timed window. The same
window with one seed (32 colonies, the batch shape of the reference's own
TSP1000 profile row) is reported beside it.
Inputs (instances, candidate rows, pheromone) are resident in HBM before the
timed region; the timed region is the continuation's iteration loop between
two stream synchronisations: the persistent kernel faco_colony_run, one
workgroup per colony running its ants and its colony update, up to 1024
iterations per launch, with HIP events around each launch on the engine's
stream.

Prints one JSON line with the metric, the roofline of that kernel (HIP-event
time per iteration, algorithmic bytes per iteration) and the CPU baseline
(the oracle's OpenMP FACO on the host cores, a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_tour(n: int, steps_per_tour: float, evals_per_tour: float, pw: int = 16,
                               ants: int = 128) -> float:
    """Bytes one ant tour must move through the persistent kernel (ant search
    and colony update): parent read + tour write (4n each), per construction
    step 16 candidate ids (4 B) and 16 trail*heuristic products (8 B), per LS
    move evaluation a candidate id, its distance and two distance lookups
    (16 B), plus the ant's share of its colony's update (trails read and
    written, heuristic read, products written: 32 B per candidate entry of
    the n x pw table, over the colony's ants)."""
    return 8.0 * n + 192.0 * steps_per_tour + 16.0 * evals_per_tour + 32.0 * n * pw / ants


def _cpu_window(xy, keys, seeds, ants, warm, steps, threads, budget_seconds):
    """Oracle FACO over colonies c = (instance c % I, seed c // I) for warm +
    steps iterations, timing the last `steps` only; colonies are cut so the
    timed window takes about `budget_seconds`."""
    from oracle import oracle
    settings = oracle.default_settings(ants=ants)
    instances = xy.shape[0]
    probe_colonies = min(len(seeds), max(threads, 1))
    probe = oracle.run_faco(xy, keys, seeds[:probe_colonies], settings, min(steps, 5), 8, threads=threads)
    per_colony = probe["search_seconds"] / probe_colonies * steps / min(steps, 5)
    colonies = int(max(1, min(len(seeds), budget_seconds / max(per_colony, 1e-9))))
    if colonies >= instances:
        colonies -= colonies % instances
    run = oracle.run_faco(xy[:min(colonies, instances)], keys[:min(colonies, instances)], seeds[:colonies], settings,
                          warm + steps, 8, threads=threads, timed_from=warm)
    tours = colonies * ants * steps
    return tours / run["search_seconds"], colonies, tours, run["search_seconds"]


def cpu_baseline(xys, keys, seeds, ants, warm, steps, budget_seconds=10.0):
    """The oracle's C++ FACO (OpenMP over colony x ant) on the host cores, on
    the GPU workload's own colonies (colony c: instance c % instances, seed
    seeds[c]), timed on the same window of the run (iterations warm+1 ..
    warm+steps); a bounded sample of the colonies, and the same window on one
    core."""
    from oracle import oracle
    threads = oracle.max_threads()
    xy = np.stack(xys).astype(np.int32)
    keys = np.asarray(keys, dtype=np.uint64)
    seeds = np.asarray(seeds, dtype=np.uint64)
    value, colonies, tours, secs = _cpu_window(xy, keys, seeds, ants, warm, steps, threads, budget_seconds)
    one, colonies1, tours1, secs1 = _cpu_window(xy, keys, seeds, ants, warm, steps, 1, budget_seconds / 5)
    return {"value": value, "unit": "ant-tours/s", "cores": threads, "kind": "port",
            "sample": f"oracle C++ FACO (OpenMP over colony x ant, -O3), the same TSP1000 instances and run: "
                      f"{colonies} colonies x {ants} ants, iterations {warm + 1}-{warm + steps} timed "
                      f"({tours} tours in {secs:.2f} s of search) after {warm} untimed; host nproc={os.cpu_count()}",
            "single_core": {"value": one, "cores": 1,
                            "sample": f"{colonies1} colonies, the same window: {tours1} tours in {secs1:.2f} s"}}


def config1_cpu_path(ext) -> dict:
    """The CPU-runnable configuration: plain FACO on one uniform TSP100
    (seed 0), 32 ants, 1000 iterations, 2-opt + Or-opt, 20-nearest
    candidates, run by the oracle's C++ FACO on one host core; the same run
    on the GPU (FixedFacoGpu) must end at the same tour."""
    from gplsaco_amd.instances import uniform_instance
    from oracle import oracle
    xy = uniform_instance(100, seed=0, instance=0)
    settings = ext.FixedFacoSettings()
    settings.ants, settings.primary_width, settings.ls_width = 32, 20, 20
    fields = ("primary_width", "backup_width", "ls_width", "beta", "retention", "p_best", "epoch_source_probability",
              "ls_evaluation_limit", "initial_ls_evaluation_limit")
    osettings = oracle.default_settings(ants=32, **{k: getattr(settings, k) for k in fields})
    cpu = oracle.run_faco(xy[None].astype(np.int32), [0], [0], osettings, 1000, 8, threads=1, or_opt=True)
    gpu = ext.FixedFacoGpu(xy, 0, settings)
    gpu.set_local_search("2opt+oropt")
    started = time.perf_counter()
    got = gpu.run_iterations(0, 1000, 8)
    gpu_seconds = time.perf_counter() - started
    return {"workload": "plain FACO, uniform TSP100 (seed 0), 32 ants, 1000 iterations, 2-opt + Or-opt, 20-NN",
            "cpu_one_core_ant_tours_per_s": 32000 / cpu["search_seconds"], "cpu_seconds": cpu["search_seconds"],
            "final_tour_length": int(cpu["costs"][0]), "gpu_final_tour_length": int(got["cost"]),
            "gpu_equals_cpu": got["tour"] == cpu["tours"][0].tolist(),
            "gpu_ant_tours_per_s": 32000 / gpu_seconds, "gpu_seconds": gpu_seconds}


def committed_counters(n: int, colonies: int, ants: int, first: int, last: int):
    """The committed rocprofv3 --pmc summary of faco_colony_run on this
    workload and window (iterations first..last), or None when none matches:
    calibrated HBM bytes per iteration (FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by what the counter reads for the calibration copy
    kernel of known size) and the issue counters."""
    path = ROOT / "profiles" / "faco_colony_run_tsp1000_counters.json"
    if not path.exists():
        return None
    data = json.loads(path.read_text())
    if (data.get("n"), data.get("colonies"), data.get("ants"), data.get("first_iteration"),
            data.get("last_iteration")) != (n, colonies, ants, first, last):
        return None
    return data


def steady_window(engine, keys, seeds, ants, warm, steps, **timing):
    """`warm` iterations from the prepared tour through the per-batch kernels
    (each colony split over two workgroups, eager launches: results
    bit-identical to the persistent kernel), then the continuation of `steps` iterations in the
    engine's automatic (production) shape; returns the continuation."""
    if not warm:
        return engine.evaluate_faco_evaluations(keys, seeds, ants * steps, **timing)
    waves = engine.launch_shape(keys.size)["waves_per_block"]
    engine.set_waves_per_block(waves, 2)
    engine.evaluate_faco_evaluations(keys, seeds, ants * warm, use_graph=False)
    engine.set_waves_per_block(0)
    return engine.evaluate_faco_evaluations(keys, seeds, ants * steps, continue_run=True, **timing)


def gp_population_rate(ext, steps: int) -> dict:
    """The GP-controlled population shape: 4 uniform TSP500 instances x 128
    random feature-spec-2 programs (512 colonies x 128 ants), `steps`
    iterations, launch shape and form (persistent controlled kernel or
    graph-replayed per-batch launches) picked by the controlled autotuner."""
    from gplsaco_amd.instances import instance_key, uniform_instance
    from gplsaco_amd.programs import random_program
    n, instances, individuals = 500, 4, 128
    rng = np.random.default_rng(0)
    programs = [{**random_program(rng, depth=5), "feature_spec_id": 2} for _ in range(individuals)]
    settings = ext.FixedFacoSettings()
    engine = ext.FacoBatchEngine(n, instances, settings, population_size=individuals)
    keys = [instance_key(n, 2, i) for i in range(instances)]
    for i, k in enumerate(keys):
        engine.register_problem(k, uniform_instance(n, seed=2, instance=i))
    ants = engine.settings.ants
    task_keys = np.array(keys, dtype=np.uint64)
    seeds = np.zeros(instances, dtype=np.uint64)
    parsed = [ext.read_program(p) for p in programs]
    tuned = engine.autotune_control(task_keys, seeds, "program", programs=parsed, iterations=8)
    raw = engine.evaluate_control_raw(task_keys, seeds, ants * steps, "program", programs=parsed, use_graph=True)
    tours = raw["result"].total_tour_evaluations
    restarts = sum(st["restarts"] for st in raw["states"])
    return {"workload": f"GP-controlled FACO, {instances} uniform TSP{n} x {individuals} random programs = "
                        f"{instances * individuals} colonies x {ants} ants, {steps} iterations",
            "ant_tours_per_s": tours / raw["result"].search_seconds, "tours": tours,
            "seconds": raw["result"].search_seconds, "restarts": restarts,
            "persistent_launches": raw["result"].persistent_launches,
            "autotune": {"best": tuned["best"], "shapes_timed": len(tuned["table"])},
            "mean_final_tour_length": float(raw["costs"].mean())}


def hard_mode_rate(ext, iterations: int) -> dict:
    """Hard mode (every relocation and 2-opt move kept inside a candidate
    graph): the headline panel (32 TSP1000 instances x 8 seeds, 128 ants,
    2-opt) with each instance's graph = primary and LS rows plus the initial
    tour, in the persistent kernel; beside it the same panel in Escape mode
    (escape gates open slot views that may add out-of-graph edges, up to 64
    per ant) and unrestricted."""
    from gplsaco_amd.instances import instance_key, uniform_instance
    n, count, seeds_per = 1000, 32, 8
    xys = [uniform_instance(n, seed=0, instance=i) for i in range(count)]
    keys = [instance_key(n, 0, i, False) for i in range(count)]
    plain = ext.FacoBatchEngine(n, count * seeds_per, ext.FixedFacoSettings())
    hard = ext.FacoBatchEngine(n, count * seeds_per, ext.FixedFacoSettings(), constraint_mode="hard")
    escape = ext.FacoBatchEngine(n, count * seeds_per, ext.FixedFacoSettings(), constraint_mode="escape")
    edges_total = 0
    for k, xy in zip(keys, xys):
        plain.register_problem(k, xy)
        data = plain.problem_data(k)
        tour = data["initial_tour"]
        edges = {tuple(sorted((int(tour[i]), int(tour[(i + 1) % n])))) for i in range(n)}
        for name in ("primary", "ls"):
            for a in range(n):
                edges.update(tuple(sorted((a, int(b)))) for b in data[name][a])
        edges_total += len(edges)
        spec = {"graph_spec_id": 1, "common_initial_tour": tour.tolist(), "edges": sorted(edges),
                "primary": data["primary"].tolist(), "backup": data["backup"].tolist(), "ls": data["ls"].tolist()}
        hard.register_graph_problem(k, xy, spec)
        escape.register_graph_problem(k, xy, spec)
    task_keys = np.array([keys[c % count] for c in range(count * seeds_per)], dtype=np.uint64)
    seeds = np.array([c // count for c in range(count * seeds_per)], dtype=np.uint64)
    ants = hard.settings.ants
    hard.evaluate_faco_evaluations(task_keys, seeds, ants * 4)      # warm-up (graph capture)
    out = hard.evaluate_faco_evaluations(task_keys, seeds, ants * iterations)
    escape.evaluate_faco_evaluations(task_keys, seeds, ants * 4)    # warm-up
    esc = escape.evaluate_faco_evaluations(task_keys, seeds, ants * iterations)
    ref = plain.evaluate_faco_evaluations(task_keys, seeds, ants * iterations)
    mean = lambda o: float(np.mean([i["cost"] for i in o["items"]]))  # noqa: E731
    return {"workload": f"Hard mode, {count} x TSP1000 x {seeds_per} seeds, {ants} ants, {iterations} iterations, "
                        f"graph = primary + LS rows + initial tour ({edges_total / count:.0f} edges per instance)",
            "ant_tours_per_s": out["total_tour_evaluations"] / out["search_seconds"],
            "unrestricted_ant_tours_per_s": ref["total_tour_evaluations"] / ref["search_seconds"],
            "mean_final_tour_length": mean(out), "unrestricted_mean_final_tour_length": mean(ref),
            "ls_evaluations_per_tour": out["completed_ls_evaluations"] / out["total_tour_evaluations"],
            "persistent_launches": out["persistent_launches"],
            "escape_mode": {"ant_tours_per_s": esc["total_tour_evaluations"] / esc["search_seconds"],
                            "mean_final_tour_length": mean(esc),
                            "ls_evaluations_per_tour": esc["completed_ls_evaluations"] / esc["total_tour_evaluations"],
                            "persistent_launches": esc["persistent_launches"],
                            "new_edges_per_tour": esc["escape_counters"]["new_edges"] / esc["total_tour_evaluations"]}}


def uneven_batch_rate(ext, rounds: int) -> dict:
    """The uneven batch: 64 instances with n uniform in [200, 2000] (32
    uniform, 32 clustered) x 128 random single-tree GP programs (depth at
    most 5, the reference's limit) x 1 seed =
    8192 colonies in one mixed-size engine (each colony with its own n and
    per-size ants: 64, 128 or 192), the whole run in one launch of the
    persistent controlled kernel, waves per workgroup picked by the
    controlled autotuner (which also times the graph-replayed per-batch
    form). The budget is `rounds` x 384 ant tours per colony (384 = a whole
    number of batches of every ant count; 167 rounds = 501 iterations of the
    128-ant colonies)."""
    from gplsaco_amd.instances import _key, _uniform_words, clustered_instance, instance_key, uniform_instance
    from gplsaco_amd.programs import random_program
    count, individuals = 64, 128
    words = _uniform_words(_key(3, 0, 0xC3), count)
    sizes = [200 + int(w % np.uint64(1801)) for w in words]
    xys, keys = [], []
    for i, n in enumerate(sizes):
        clustered = i % 2 == 1
        xys.append((clustered_instance if clustered else uniform_instance)(n, seed=3, instance=i))
        keys.append(instance_key(n, 3, i, clustered))
    rng = np.random.default_rng(3)
    programs = [ext.read_program({**random_program(rng, depth=5), "feature_spec_id": 2}) for _ in range(individuals)]
    engine = ext.MixedFacoBatchEngine(max(sizes), count, ext.FixedFacoSettings(), population_size=individuals)
    # preparation (candidate rows, nearest-neighbour tour, initial 2-opt) of
    # the 64 instances on the device, and the same on the host for comparison
    started = time.perf_counter()
    engine.register_problems_device(keys, xys)
    prep = time.perf_counter() - started
    host_engine = ext.MixedFacoBatchEngine(max(sizes), count, ext.FixedFacoSettings())
    started = time.perf_counter()
    for k, xy in zip(keys, xys):
        host_engine.register_problem(k, xy)
    host_prep = time.perf_counter() - started
    same = all(np.array_equal(engine.problem_data(k)["initial_tour"], host_engine.problem_data(k)["initial_tour"])
               for k in keys)
    del host_engine
    task_keys = np.array(keys, dtype=np.uint64)
    seeds = np.zeros(count, dtype=np.uint64)
    # the controlled autotuner on this run; the run itself is one persistent
    # launch at the best wave count
    tuned = engine.autotune_control(task_keys, seeds, "program", programs=programs, iterations=2)
    best = tuned["best"]
    engine.set_waves_per_block(best["waves"] | (0 if best["neighbour_rows_in_lds"] else 0x100), 1)
    limit = 384 * rounds
    raw = engine.evaluate_control_raw(task_keys, seeds, limit, "program", programs=programs, use_graph=True)
    res = raw["result"]
    per_batch = [r for r in tuned["table"] if not r["persistent"]]
    ants = [64 * math.ceil(math.sqrt(n) / 16) if n > 256 else 64 for n in sizes]
    return {"workload": f"GP-controlled FACO, {count} instances n in [200,2000] (half clustered) x {individuals} "
                        f"random programs = {count * individuals} colonies, {limit} ant tours per colony",
            "sizes": {"min": min(sizes), "max": max(sizes), "mean": float(np.mean(sizes))},
            "ant_tours_per_s": res.total_tour_evaluations / res.search_seconds,
            "tours": res.total_tour_evaluations, "seconds": res.search_seconds,
            "launched_iterations": res.launched_batches, "persistent_launches": res.persistent_launches,
            "iterations_per_colony": {str(a): limit // a for a in (64, 128, 192)},
            "preparation_seconds": prep,
            "preparation": {"device_seconds": prep, "host_seconds": host_prep, "host_threads": os.cpu_count(),
                            "identical_initial_tours": same},
            "restarts": sum(st["restarts"] for st in raw["states"]),
            "autotune": {"best": best, "shapes_timed": len(tuned["table"]),
                         "best_per_batch_form": min(per_batch, key=lambda r: r["ms_per_iteration"]) if per_batch else None},
            "ants_per_colony": {"64": ants.count(64), "128": ants.count(128), "192": ants.count(192)}}


def full_pipeline(ext, rounds: int) -> dict:
    """The full pipeline at TSP1000 and TSP2000: 10 + 10 instances (half
    clustered) x 128 random conditional three-tree GP controllers in one
    mixed-size engine, the learned edge prior on (k = 20 candidates),
    launch shape and form picked by the controlled autotuner, `rounds` x 384
    ant tours per colony (500 rounds = 1000 iterations of the TSP2000
    colonies' 192 ants, 1500 of the TSP1000 colonies' 128); plain FACO (same
    k, no prior, no control) on the same instances with 8 seeds and the same
    budget for the gap."""
    from gplsaco_amd import prior as P
    from gplsaco_amd.programs import random_controller
    sizes = [1000] * 10 + [2000] * 10
    settings = ext.FixedFacoSettings()
    settings.primary_width = 20
    keys, xys = [], []
    for i, n in enumerate(sizes):
        key, xy = P.synthetic_instance(n, i % 2 == 1, i, 5)
        keys.append(key)
        xys.append(xy)
    rng = np.random.default_rng(5)
    controllers = [ext.read_controller(random_controller(rng, f"c{i}")) for i in range(128)]
    limit = 384 * rounds
    engine = ext.MixedFacoBatchEngine(2000, len(sizes), settings, population_size=len(controllers))
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    net = P.EdgePrior.load()
    for k in keys:
        engine.apply_edge_prior(net, k)
    task_keys = np.array(keys, dtype=np.uint64)
    seeds = np.zeros(len(sizes), dtype=np.uint64)
    tuned = engine.autotune_control(task_keys, seeds, "controller", controllers=controllers, iterations=2)
    raw = engine.evaluate_control_raw(task_keys, seeds, limit, "controller", controllers=controllers, use_graph=True)
    res = raw["result"]
    costs = raw["costs"].reshape(len(controllers), len(sizes))
    plain_engine = ext.MixedFacoBatchEngine(2000, len(sizes) * 8, settings)
    for k, xy in zip(keys, xys):
        plain_engine.register_problem(k, xy)
    plain = plain_engine.evaluate_faco_evaluations(np.tile(task_keys, 8), np.repeat(np.arange(8, dtype=np.uint64), len(sizes)),
                                                   limit)
    plain_costs = np.array([it["cost"] for it in plain["items"]]).reshape(8, len(sizes))
    out = {"workload": f"GP-controlled FACO with the learned prior, 10 x TSP1000 + 10 x TSP2000 (half clustered) x "
                       f"{len(controllers)} random conditional three-tree controllers = {len(sizes) * len(controllers)} "
                       f"colonies, k = 20, {limit} ant tours per colony",
           "ant_tours_per_s": res.total_tour_evaluations / res.search_seconds, "tours": res.total_tour_evaluations,
           "seconds": res.search_seconds, "restarts": sum(st["restarts"] for st in raw["states"]),
           "autotune": tuned["best"], "persistent_launches": res.persistent_launches,
           "iterations_per_colony": {"tsp1000": limit // 128, "tsp2000": limit // 192},
           "plain_faco_ant_tours_per_s": plain["total_tour_evaluations"] / plain["search_seconds"]}
    for n in (1000, 2000):
        cols = [i for i, m in enumerate(sizes) if m == n]
        controlled, base = float(costs[:, cols].mean()), float(plain_costs[:, cols].mean())
        best = float(costs[:, cols].mean(axis=1).min())
        out[f"tsp{n}"] = {"mean_final_tour_length": controlled, "plain_faco_mean": base,
                          "gap_vs_plain": (controlled - base) / base,
                          "best_controller_gap_vs_plain": (best - base) / base}
    return out


def learned_prior_legs(ext, budgets) -> dict:
    """Learned edge prior: (1) the committed trained network against plain
    FACO at equal FE on unseen TSP1000 instances (16: half uniform, half
    clustered, x 8 seeds, k = 20 candidates) for each iteration budget;
    (2) training throughput of the prior's forward+backward+Adam kernels on
    those instances' candidate edges (labels: plain FACO's best tours)."""
    from gplsaco_amd import prior as P
    if not P.DEFAULT_WEIGHTS.exists():
        return {"skipped": "no trained weights"}
    net = P.EdgePrior.load()
    out = {"weights": P.DEFAULT_WEIGHTS.name, "training": {k: v for k, v in net.info.items() if k != "sizes"}}
    for its in budgets:
        out[f"budget_{its}_iterations"] = P.equal_budget_comparison(net, iterations=its, instances=16, seeds=8)
    # training throughput: 8 TSP1000 instances' edges, replicated to 5.1e6 edges
    settings = ext.FixedFacoSettings()
    settings.primary_width = 20
    engine = ext.FacoBatchEngine(1000, 8, settings)
    keys = []
    for i in range(8):
        key, xy = P.synthetic_instance(1000, i % 2 == 1, i, 104729)
        engine.register_problem(key, xy)
        keys.append(key)
    res = engine.evaluate_faco_evaluations(np.array(keys, np.uint64), np.zeros(8, np.uint64), engine.settings.ants * 20)
    feats = [engine.edge_features(k) for k in keys]
    labels = [P.tour_labels(engine.problem_data(k)["primary"], it["tour"]) for k, it in zip(keys, res["items"])]
    f = np.tile(np.concatenate(feats), (32, 1))
    y = np.tile(np.concatenate(labels), 32)
    scratch = P.EdgePrior(seed=9)
    scratch.train(f[:65536], y[:65536], epochs=1)
    started = time.perf_counter()
    scratch.train(f, y, batch=8192, epochs=2)
    seconds = time.perf_counter() - started
    flops = 3 * 2 * (8 * 64 + 64 * 64 + 64) * f.shape[0] * 2
    started = time.perf_counter()
    scratch.logits(f)
    infer_seconds = time.perf_counter() - started
    out["training_throughput"] = {
        "edges": int(f.shape[0]), "epochs": 2, "batch": 8192, "seconds": seconds,
        "edges_per_s": 2 * f.shape[0] / seconds, "algorithmic_tflops": flops / seconds / 1e12,
        "bf16_peak_tflops": 2500.0, "note": "host shuffle and upload per epoch included",
        "inference_edges_per_s_including_copies": f.shape[0] / infer_seconds}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed iterations: the last ones of the run")
    ap.add_argument("--warmup", type=int, default=5, help="untimed iterations before them (at least)")
    ap.add_argument("--run-length", type=int, default=100,
                    help="iterations of the run whose last --steps are timed (warm-up = run length - steps, "
                         "at least --warmup)")
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--population-steps", type=int, default=1000,
                    help="iterations of the GP-controlled population measurement (0 skips it)")
    ap.add_argument("--hard-iterations", type=int, default=20,
                    help="iterations of the Hard-mode leg (0 skips it)")
    ap.add_argument("--uneven-rounds", type=int, default=167,
                    help="budget of the uneven-size GP batch in units of 384 ant tours per colony (0 skips it)")
    ap.add_argument("--pipeline-rounds", type=int, default=500,
                    help="budget of the full-pipeline leg in units of 384 ant tours per colony (0 skips it)")
    ap.add_argument("--prior-budgets", default="100,300",
                    help="iteration budgets of the learned-prior comparison (empty skips it)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the bench, run the calibration copy kernel over this many MiB (counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine's colonies are independent and every run here uses a "
              "single MI355X; --gpus above 1 is not measured", file=sys.stderr)
        return 2

    from gplsaco_amd import gp_faco_ext as ext
    from gplsaco_amd.instances import instance_key, uniform_instance

    n, instances = args.n, args.instances
    colonies = instances * args.seeds
    xys = [uniform_instance(n, seed=0, instance=i) for i in range(instances)]
    keys = [instance_key(n, 0, i) for i in range(instances)]
    settings = ext.FixedFacoSettings()
    engine = ext.FacoBatchEngine(n, colonies, settings)
    ants = engine.settings.ants
    prep_started = time.perf_counter()
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    prep_seconds = time.perf_counter() - prep_started
    task_keys = np.array([keys[c % instances] for c in range(colonies)], dtype=np.uint64)
    task_seeds = np.array([c // instances for c in range(colonies)], dtype=np.uint64)
    print(f"# prepared {instances} x TSP{n} in {prep_seconds:.2f} s; shape {engine.launch_shape()}",
          file=sys.stderr, flush=True)
    warm = max(args.warmup, args.run_length - args.steps)
    first, last = warm + 1, warm + args.steps
    # The timed window: the continuation of the warm-up, HIP events around
    # every launch of the production kernel (the persistent faco_colony_run:
    # ant search and colony update, up to 1024 iterations per launch) on the
    # engine's stream; its wall clock is the bench value.
    timed = steady_window(engine, task_keys, task_seeds, ants, warm, args.steps, time_kernels=True)
    print(f"# timed iterations {first}-{last} in {timed['search_seconds']:.3f} s", file=sys.stderr, flush=True)
    tours = timed["total_tour_evaluations"]
    seconds = timed["search_seconds"]
    value = tours / seconds
    steps_per_tour = timed["completed_construction_steps"] / tours
    evals_per_tour = timed["completed_ls_evaluations"] / tours
    launches = -(-args.steps // 1024)
    iteration_ms = (timed["kernel_milliseconds"]["construction_and_search"]
                    + timed["kernel_milliseconds"]["reduction_and_pheromone"]) / args.steps
    # The last iterations of the window again with the clock-instrumented
    # kernel variant (per-batch launches): per-ant phases.
    phase_steps = min(args.steps, 10)
    phases = steady_window(engine, task_keys, task_seeds, ants, last - phase_steps, phase_steps, use_graph=False,
                           profile_kernels=True)
    bytes_per_tour = algorithmic_bytes_per_tour(n, steps_per_tour, evals_per_tour, engine.settings.primary_width, ants)
    tours_per_iteration = colonies * ants
    achieved = bytes_per_tour * tours_per_iteration / (iteration_ms * 1e-3) / 1e9
    mean_cost = float(np.mean([item["cost"] for item in timed["items"]]))
    counters = committed_counters(n, colonies, ants, first, last)

    # The same window with Or-opt added to the local search.
    engine.set_local_search("2opt+oropt")
    oropt = steady_window(engine, task_keys, task_seeds, ants, warm, args.steps)
    engine.set_local_search("2opt")
    oropt_value = oropt["total_tour_evaluations"] / oropt["search_seconds"]
    oropt_cost = float(np.mean([item["cost"] for item in oropt["items"]]))

    # The reference's profile shape: 32 colonies (one seed per instance),
    # the same window.
    small = ext.FacoBatchEngine(n, instances, settings)
    for k, xy in zip(keys, xys):
        small.register_problem(k, xy)
    small_keys, small_seeds = np.array(keys, dtype=np.uint64), np.zeros(instances, dtype=np.uint64)
    small_run = steady_window(small, small_keys, small_seeds, ants, warm, args.steps)
    small_value = small_run["total_tour_evaluations"] / small_run["search_seconds"]

    line = {
        "metric": "ant tours/s (construct+2-opt/Or-opt) at TSP1000; gap vs plain FACO at equal FE",
        "value": value,
        "unit": "ant-tours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": seconds / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain FACO (no GP, fixed MNE 8), {instances} uniform random EUC_2D TSP{n} instances x "
                        f"{args.seeds} seeds = {colonies} colonies, {ants} ants each, 16/64/20 candidates, "
                        f"checklist 2-opt; iterations {first}-{last} of a {last}-iteration run timed",
            "n": n, "colonies": colonies, "ants": ants, "mne": 8, "iterations": args.steps,
            "timed_iterations": [first, last], "warmup_iterations": warm,
            "value_at_32_colonies": small_value,
            "value_with_or_opt": oropt_value, "mean_final_tour_length_with_or_opt": oropt_cost,
            "ls_evaluations_per_tour_with_or_opt": oropt["completed_ls_evaluations"] / oropt["total_tour_evaluations"],
            "construction_steps_per_tour": steps_per_tour, "ls_evaluations_per_tour": evals_per_tour,
            "mean_final_tour_length": mean_cost, "launch_shape": engine.launch_shape(colonies),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "faco_colony_run",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": counters["bytes_per_iteration"] if counters else None,
            "limiter": "instruction issue (VALU and SALU), not HBM: see issue",
            "issue": counters["issue"] if counters else None,
            "per": "iteration (launches run up to 1024 iterations each; achieved and traffic are per iteration)",
            "bytes_per_tour": bytes_per_tour,
            "tours_per_iteration": tours_per_iteration,
            "iteration_ms": iteration_ms,
            "launches": launches,
            "mean_launch_ms": iteration_ms * args.steps / launches,
            "mean_ant_phase_ms_per_step": {k: v / phase_steps for k, v in phases["ant_phase_milliseconds"].items()},
            "mean_ant_subphase_megaclocks_per_step": {k: v / phase_steps
                                                      for k, v in phases["ant_subphase_megaclocks"].items()},
            "mean_ant_hold_ms_per_step": phases["ant_hold_milliseconds"] / phase_steps,
            "wave_utilisation": phases["wave_utilisation"] / phase_steps,
        },
        "cpu_baseline": None,
    }
    if args.population_steps:
        line["config"]["gp_population"] = gp_population_rate(ext, args.population_steps)
    if args.hard_iterations:
        line["config"]["hard_mode"] = hard_mode_rate(ext, args.hard_iterations)
    if args.uneven_rounds:
        line["config"]["uneven_batch"] = uneven_batch_rate(ext, args.uneven_rounds)
    if args.pipeline_rounds:
        line["config"]["full_pipeline"] = full_pipeline(ext, args.pipeline_rounds)
    if args.prior_budgets:
        line["config"]["learned_prior"] = learned_prior_legs(ext, [int(b) for b in args.prior_budgets.split(",")])
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(xys, keys, task_seeds, ants, warm, args.steps)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
        line["gpu_over_one_core"] = value / line["cpu_baseline"]["single_core"]["value"]
        line["config"]["config1_cpu_path"] = config1_cpu_path(ext)
    if args.copy_mib:
        copy_ms = ext.diagnostic_copy(args.copy_mib << 20, 4)
        line["calibration_copy"] = {"bytes_read": args.copy_mib << 20, "bytes_written": args.copy_mib << 20,
                                    "ms": copy_ms, "GB/s": 2 * (args.copy_mib << 20) / copy_ms / 1e6}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
