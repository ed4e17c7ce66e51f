"""GPU parity of the HIP engine against the CPU oracle (bit-exact).

The oracle itself is pinned to the reference's CPU core and golden vectors by
the CPU tests (test_oracle_*.py). Every comparison here is exact equality:
tours, positions, checklists, integer costs, work counters, trails.
"""
import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu

U64 = (1 << 64) - 1


def rows_by_distance(xy, width):
    from oracle import oracle as O
    M = O.distance_matrix(xy.astype(np.int64))
    n = xy.shape[0]
    order = np.lexsort((np.broadcast_to(np.arange(n), (n, n)), M), axis=1)
    rows = np.zeros((n, width), dtype=np.uint32)
    for a in range(n):
        rows[a] = [b for b in order[a] if b != a][:width]
    return rows


def test_device_distances_exact(ext, oracle_mod):
    rng = np.random.default_rng(11)
    pairs = rng.integers(0, 32768, size=(20000, 4))
    edge = np.array([[0, 0, 32767, 32767], [0, 0, 0, 0], [5, 5, 5, 5], [0, 0, 3, 4], [0, 0, 1, 1],
                     [0, 0, 2, 2], [0, 0, 7, 24], [32767, 0, 0, 32767], [100, 100, 103, 104]])
    # perfect squares s = k^2
    k = rng.integers(1, 23000, size=2000)
    # the rounding boundaries: s = R^2 + R rounds down to R (dx = m^2, dy = m,
    # R = m^2) and s = R^2 + R + 1 up to R + 1 (dx = m^2 - 1, dy = m, R = m^2 - 1)
    m = np.arange(1, 182)
    zero = 0 * m
    boundary = np.concatenate([np.stack([zero, zero, m * m, m], 1), np.stack([zero, zero, m * m - 1, m], 1),
                               np.stack([32767 - m * m, 32767 - m, zero + 32767, zero + 32767], 1)])
    pairs = np.concatenate([pairs, edge, np.stack([0 * k, 0 * k, k, 0 * k], 1), boundary]).astype(np.int32)
    announce("distances", pairs.shape[0])
    got = ext.diagnostic_distances(pairs)
    d = pairs[:, :2].astype(np.int64) - pairs[:, 2:].astype(np.int64)
    s = (d * d).sum(1)
    want = np.floor(np.sqrt(s.astype(np.float64)) + 0.5).astype(np.int64)
    np.testing.assert_array_equal(got, want)
    for i in range(0, pairs.shape[0], 97):
        assert oracle_mod.euc2d(*map(int, pairs[i])) == got[i]


def test_device_philox_matches_oracle(ext, oracle_mod):
    for key, batch, ant, step in [(0, 0, 0, 0), (U64, 7, 3, 1), (0x1234567890ABCDEF, 99, 127, 0xFFFFFFFF),
                                  (42, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFE)]:
        announce("philox", key, batch, ant, step)
        words = ext.diagnostic_philox(key, batch, ant, step, 12)
        want = []
        ctr = [step, 0, ant, batch]
        for _ in range(3):
            want.extend(oracle_mod.philox(ctr, [key & 0xFFFFFFFF, key >> 32]).tolist())
            value = (ctr[0] | ctr[1] << 32 | ctr[2] << 64 | ctr[3] << 96) + 1
            ctr = [(value >> (32 * i)) & 0xFFFFFFFF for i in range(4)]
        np.testing.assert_array_equal(words, np.array(want, dtype=np.uint32))


CAPS = [0, 1, 7, 127, U64]


@pytest.mark.parametrize("n", [3, 4, 5, 7, 8, 17, 100, 500, 1000, 1024, 1025, 1500, 2048, 2049])
@pytest.mark.parametrize("metric", ["uniform", "grid", "repeated"])
@pytest.mark.parametrize("ls_mode", ["2opt", "2opt+oropt"])
def test_explicit_construct_ls_matches_oracle(ext, oracle_mod, n, metric, ls_mode):
    rng = np.random.default_rng(83741 + n + len(metric))
    if metric == "uniform":
        xy = rng.integers(0, 10001, size=(n, 2))
    elif metric == "grid":
        xy = rng.integers(0, 13, size=(n, 2))
    else:
        xy = rng.integers(0, 10001, size=(n, 2))
        xy[3::3] = xy[0]
    width = min(20, n - 1)
    rows = rows_by_distance(xy, width)
    ants = 8
    for cap in CAPS:
        parents = np.stack([rng.permutation(n) for _ in range(ants)]).astype(np.uint32)
        visits = np.stack([rng.permutation(n) for _ in range(ants)]).astype(np.uint32)
        targets = np.array([2, 4, 8, 16] * 2, dtype=np.uint32)
        waves = [1, 2, 4, 8][cap % 4 if cap < 8 else 3]
        chunk = min(8, [waves, 2 * waves, 8][CAPS.index(cap) % 3])
        announce("construct_ls", n, metric, cap, waves, chunk, ls_mode)
        got = ext.diagnostic_construct_ls(xy.astype(np.float64), rows, parents, visits, targets, cap, waves, chunk,
                                          local_search=ls_mode)
        for a in range(ants):
            want = oracle_mod.construct_ls(xy, rows, parents[a], visits[a], int(targets[a]), cap,
                                           or_opt=ls_mode != "2opt")
            info = got["info"][a]
            c = want["counters"]
            assert np.array_equal(got["construction_tours"][a], want["construction_tour"]), (cap, a)
            assert np.array_equal(got["tours"][a], want["tour"]), (cap, a)
            assert np.array_equal(got["positions"][a], want["positions"]), (cap, a)
            assert info["checklist_size"] == c["checklist_size"], (cap, a)
            assert np.array_equal(got["checklists"][a][: c["checklist_size"]], want["checklist"]), (cap, a)
            for name in ("mne", "steps", "nonidentity_relocations", "processed_nodes", "candidate_checks",
                         "move_evaluations", "accepted_moves", "reactivations", "evaluation_limit_reached"):
                assert int(info[name]) == c[name], (name, cap, a)
            assert int(info["construction_cost"]) == want["construction_cost"]
            assert int(info["final_cost"]) == want["final_cost"]


def _settings(ext, **kw):
    s = ext.FixedFacoSettings()
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _oracle_settings(oracle_mod, s):
    return oracle_mod.default_settings(**{k: getattr(s, k) for k in (
        "ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
        "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")})


VARIANTS = [  # (primary, backup, retention, source probability): tests/cpp/test_fixed_faco.cpp:41-47
    (16, 64, 0.25, 0.0), (2, 1, 0.5, 0.01), (16, 64, 0.75, 1.0), (2, 1, 0.9, 0.3)]


@pytest.mark.parametrize("n", [7, 17, 100, 500])
@pytest.mark.parametrize("variant", range(4))
def test_fixed_faco_trace_matches_oracle(ext, oracle_mod, n, variant):
    pw, bw, rho, src = VARIANTS[variant]
    rng = np.random.default_rng(4703 + n)
    xy = rng.integers(0, 10001, size=(n, 2))
    if variant == 3:
        xy[3::3] = xy[0]
    s = _settings(ext, ants=8, primary_width=pw, backup_width=bw, retention=rho, epoch_source_probability=src,
                  ls_evaluation_limit=5000)
    key, seed, batches = n + 8999, 17, 8
    announce("fixed", n, variant)
    engine = ext.FixedFacoGpu(xy.astype(np.float64), key, s)
    got = engine.run_iterations(seed, batches, 16, trace=True)
    want = oracle_mod.run_faco(xy[None].astype(np.int32), [key], [seed], _oracle_settings(oracle_mod, s), batches,
                               16, threads=1, trace=True)
    tr = got["trace"]
    np.testing.assert_array_equal(tr["parent"], want["trace_parent"])
    np.testing.assert_array_equal(tr["trails"], want["trace_trails"])
    np.testing.assert_array_equal(tr["costs"], want["trace_costs"])
    assert got["tour"] == want["tours"][0].tolist()
    assert got["cost"] == float(want["costs"][0])
    assert got["construction_steps"] == int(want["work"][0, 0])
    assert got["ls_evaluations"] == int(want["work"][0, 1])
    # a second run with another seed, then the original again, reproduces it
    engine.run_iterations(29, 2, 2)
    again = engine.run_iterations(seed, batches, 16)
    assert again["tour"] == got["tour"] and again["cost"] == got["cost"]
    zero = engine.run_iterations(seed, 0, 8)
    assert zero["cost"] == got["initial_cost"]


def _panel(n, instances, clustered=False):
    gen = clustered_instance if clustered else uniform_instance
    xys = [gen(n, seed=3, instance=i) for i in range(instances)]
    keys = [1000 + i for i in range(instances)]
    return xys, keys


@pytest.mark.parametrize("n,instances,seeds,batches,clustered", [
    (100, 4, 2, 10, False), (200, 2, 3, 20, True), (1000, 2, 1, 4, False), (1300, 2, 1, 3, True)])
def test_batch_engine_matches_oracle(ext, oracle_mod, n, instances, seeds, batches, clustered):
    xys, keys = _panel(n, instances, clustered)
    s = _settings(ext, ants=32 if n < 1000 else 64)
    colonies = instances * seeds
    task_keys = np.array([keys[c % instances] for c in range(colonies)], dtype=np.uint64)
    task_seeds = np.array([5 + c // instances for c in range(colonies)], dtype=np.uint64)
    engine = ext.FacoBatchEngine(n, colonies, s)
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    announce("batch", n, colonies, batches)
    got = engine.evaluate_faco_evaluations(task_keys, task_seeds, s.ants * batches, 8, use_graph=batches >= 16)
    want = oracle_mod.run_faco(np.stack(xys).astype(np.int32), keys, task_seeds, _oracle_settings(oracle_mod, s),
                               batches, 8, threads=0)
    for c in range(colonies):
        assert got["items"][c]["tour"] == want["tours"][c].tolist(), c
        assert got["items"][c]["cost"] == float(want["costs"][c]), c
    np.testing.assert_array_equal(got["colony_work"], want["work"])
    assert got["total_tour_evaluations"] == colonies * s.ants * batches
    assert got["completed_construction_steps"] == int(want["work"][:, 0].sum())


@pytest.mark.parametrize("n,instances,seeds,batches", [(60, 2, 2, 12), (300, 2, 1, 6), (1000, 1, 2, 3)])
def test_batch_engine_or_opt_matches_oracle(ext, oracle_mod, n, instances, seeds, batches):
    xys, keys = _panel(n, instances, clustered=n == 300)
    s = _settings(ext, ants=32 if n < 1000 else 64)
    colonies = instances * seeds
    task_keys = np.array([keys[c % instances] for c in range(colonies)], dtype=np.uint64)
    task_seeds = np.array([9 + c // instances for c in range(colonies)], dtype=np.uint64)
    engine = ext.FacoBatchEngine(n, colonies, s)
    engine.set_local_search("2opt+oropt")
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    announce("batch-oropt", n, colonies, batches)
    got = engine.evaluate_faco_evaluations(task_keys, task_seeds, s.ants * batches, 8, use_graph=False)
    want = oracle_mod.run_faco(np.stack(xys).astype(np.int32), keys, task_seeds, _oracle_settings(oracle_mod, s),
                               batches, 8, threads=0, or_opt=True)
    for c in range(colonies):
        assert got["items"][c]["tour"] == want["tours"][c].tolist(), c
    np.testing.assert_array_equal(got["colony_work"], want["work"])
    plain = ext.FacoBatchEngine(n, colonies, s)
    for k, xy in zip(keys, xys):
        plain.register_problem(k, xy)
    base = plain.evaluate_faco_evaluations(task_keys, task_seeds, s.ants * batches, 8)
    assert base["completed_ls_evaluations"] < got["completed_ls_evaluations"]


def test_graph_replay_and_layouts_are_identical(ext):
    n, instances = 300, 3
    xys, keys = _panel(n, instances)
    s = _settings(ext, ants=64)
    engine = ext.FacoBatchEngine(n, instances, s)
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
    task_keys = np.array(keys, dtype=np.uint64)
    seeds = np.arange(instances, dtype=np.uint64)
    announce("graph-eager", n)
    base = engine.evaluate_faco_evaluations(task_keys, seeds, 64 * 40, use_graph=False)
    for rep in range(3):
        again = engine.evaluate_faco_evaluations(task_keys, seeds, 64 * 40, use_graph=True)
        assert [i["tour"] for i in again["items"]] == [i["tour"] for i in base["items"]], rep
        np.testing.assert_array_equal(again["colony_work"], base["colony_work"])
    first = None
    for waves, bpc in ((1, 0), (2, 0), (4, 0), (8, 0), (16, 0), (16, 1), (8, 2), (4, 16)):
        engine.set_waves_per_block(waves, bpc)
        announce("layout", waves, bpc, engine.launch_shape())
        other = engine.evaluate_faco_evaluations(task_keys, seeds, 64 * 20, use_graph=True)
        if first is None:
            first = other
            continue
        assert [i["tour"] for i in other["items"]] == [i["tour"] for i in first["items"]], waves
        np.testing.assert_array_equal(other["colony_work"], first["colony_work"])


def test_tsp1000_full_size_properties(ext, oracle_mod):
    n, instances = 1000, 4
    xys, keys = _panel(n, instances)
    s = _settings(ext)
    engine = ext.FacoBatchEngine(n, instances, s)
    initial = []
    for k, xy in zip(keys, xys):
        engine.register_problem(k, xy)
        initial.append(engine.problem_data(k)["initial_cost"])
    announce("tsp1000", instances)
    ants = engine.settings.ants
    assert ants == 128
    got = engine.evaluate_faco_evaluations(np.array(keys, dtype=np.uint64), np.arange(instances, dtype=np.uint64),
                                           ants * 20, profile_kernels=True)
    assert got["completed_batches"] == 20
    for c, item in enumerate(got["items"]):
        tour = np.array(item["tour"])
        assert np.array_equal(np.sort(tour), np.arange(n))
        M = oracle_mod.distance_matrix(xys[c].astype(np.int64))
        assert int(M[tour, np.roll(tour, -1)].sum()) == item["cost"]
        assert item["cost"] <= initial[c]
    assert got["kernel_milliseconds"]["construction_and_search"] > 0


def test_score_gpu_matches_score_cpu(ext):
    rng = np.random.default_rng(6677)
    features = rng.random((12, 7, 32), dtype=np.float32)
    masks = rng.integers(1, 0xFFFFFFFF, size=7, dtype=np.uint32)
    masks[:3] = [1, 1 << 31, 0xFFFFFFFF]
    from gplsaco_amd.programs import random_program
    for i in range(200):
        program = random_program(rng, depth=5)
        a = ext.score_cpu(program, features, masks)
        b = ext.score_cuda(program, features, masks)
        np.testing.assert_array_equal(a["scores"].view(np.uint32), b["scores"].view(np.uint32))
        np.testing.assert_array_equal(a["actions"], b["actions"])


def test_device_start_nodes_match_oracle(ext, oracle_mod):
    counts = np.array([3, 7, 1000, 1024, 10000, 65535, (1 << 31) + 1, (1 << 31) + 7, 0xFFFFFFFF, 3000000001],
                      dtype=np.uint32)
    for key, batch, ant in [(0, 0, 0), (0x9E3779B97F4A7C15, 5, 77), (2**64 - 1, 2**32 - 1, 2**32 - 1)]:
        announce("start", key, batch, ant)
        dev = ext.diagnostic_start_nodes(key, batch, ant, counts, on_device=True)
        want = np.array([oracle_mod.start_node(key, batch, ant, int(c)) for c in counts], dtype=np.uint32)
        np.testing.assert_array_equal(dev, want)


@pytest.mark.parametrize("n,pw", [(300, 16), (120, 8), (200, 2)])
def test_fast_roulette_equals_sequential_sums(ext, oracle_mod, n, pw):
    """The roulette's tree-order fast path never changes a selection: with the
    margin at +inf every step uses the exact left-to-right sums, and both runs
    equal the oracle (which sums left to right)."""
    xys, keys = _panel(n, 2, clustered=True)
    s = _settings(ext, ants=32, primary_width=pw, backup_width=min(64, n - 1 - pw))
    colonies = 4
    task_keys = np.array([keys[c % 2] for c in range(colonies)], dtype=np.uint64)
    task_seeds = np.array([11 + c // 2 for c in range(colonies)], dtype=np.uint64)
    runs = []
    for margin in (None, float("inf")):
        engine = ext.FacoBatchEngine(n, colonies, s)
        for k, xy in zip(keys, xys):
            engine.register_problem(k, xy)
        if margin is not None:
            engine.set_roulette_margin(margin)
        announce("roulette", n, pw, margin)
        runs.append(engine.evaluate_faco_evaluations(task_keys, task_seeds, s.ants * 12, 8, use_graph=False))
    want = oracle_mod.run_faco(np.stack(xys).astype(np.int32), keys, task_seeds, _oracle_settings(oracle_mod, s), 12, 8)
    for run in runs:
        for c in range(colonies):
            assert run["items"][c]["tour"] == want["tours"][c].tolist(), c
        np.testing.assert_array_equal(run["colony_work"], want["work"])
    with pytest.raises(ValueError):
        ext.FacoBatchEngine(n, colonies, s).set_roulette_margin(0.0)
