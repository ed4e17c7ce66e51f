"""Hard mode on the device (constraint_mode "hard", register_graph_problem)
against the oracle's Hard runs (oracle.run_faco(..., graphs=...), itself
pinned in tests/test_oracle_hard.py): the same tours, costs and work counters
bit for bit, every tour inside its graph, on the linked-list (n <= 2048) and
the array (n > 2048) construction paths, per batch and in the persistent
kernel (one workgroup per colony running its batches and colony updates). The graph is the reference's usual
candidate graph: the primary and LS rows plus the initial tour
(SparseUndirectedGraph::from_candidates, cpp/src/sparse_graph.cpp:27-40),
optionally with a few random extra edges. Controlled runs (a GP program,
a rule baseline) in Hard mode equal the oracle's controlled Hard runs too
(tours, work, every action, restarts).
"""
import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def _cycle(tour):
    return {tuple(sorted((int(tour[i]), int(tour[(i + 1) % len(tour)])))) for i in range(len(tour))}


def _spec(data, extra_edges):
    n = data["primary"].shape[0]
    edges = set(_cycle(data["initial_tour"])) | set(extra_edges)
    for name in ("primary", "ls"):
        for a in range(n):
            for b in data[name][a]:
                if b < n:
                    edges.add(tuple(sorted((a, int(b)))))
    return {"graph_spec_id": 1, "common_initial_tour": data["initial_tour"].tolist(),
            "edges": [list(e) for e in sorted(edges)], "primary": data["primary"].tolist(),
            "backup": data["backup"].tolist(), "ls": data["ls"].tolist()}, edges


@pytest.mark.parametrize("n,extra,persistent", [(200, 0, False), (300, 150, False), (1100, 0, False), (2100, 0, False),
                                                (300, 150, True), (1100, 0, True), (2100, 0, True)])
def test_hard_mode_matches_the_oracle_and_stays_in_the_graph(ext, oracle_mod, n, extra, persistent):
    O = oracle_mod
    count, batches = 2, 5
    xys = [(clustered_instance if i % 2 else uniform_instance)(n, seed=41, instance=i) for i in range(count)]
    keys = [300 + i for i in range(count)]
    s = ext.FixedFacoSettings()
    s.ants = 32
    plain = ext.FacoBatchEngine(n, 4, s)
    hard = ext.FacoBatchEngine(n, 4, s, constraint_mode="hard")
    rng = np.random.default_rng(n)
    graphs, members = [], []
    for k, xy in zip(keys, xys):
        plain.register_problem(k, xy)
        pairs = {tuple(sorted(map(int, rng.choice(n, 2, replace=False)))) for _ in range(extra)}
        spec, edges = _spec(plain.problem_data(k), pairs)
        hard.register_graph_problem(k, xy, spec)
        graphs.append(O.csr(sorted(edges), n))
        members.append(edges)
    task_keys = np.array([keys[c % count] for c in range(4)], dtype=np.uint64)
    seeds = np.array([1, 1, 2, 2], dtype=np.uint64)
    if persistent:
        # one workgroup per colony: the persistent kernel runs every batch
        hard.set_waves_per_block(hard.launch_shape(4)["waves_per_block"], 1)
        assert hard.launch_shape(4)["blocks_per_colony"] == 1
    announce("hard", n, extra, persistent)
    got = hard.evaluate_faco_evaluations(task_keys, seeds, s.ants * batches)
    assert (got["persistent_launches"] > 0) == persistent, got["persistent_launches"]
    ref = plain.evaluate_faco_evaluations(task_keys, seeds, s.ants * batches)
    settings = O.default_settings(**{f: getattr(s, f) for f in (
        "ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
        "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")})
    want = O.run_faco(np.stack(xys).astype(np.int32), keys, seeds, settings, batches, 8, graphs=graphs)
    for c in range(4):
        tour = got["items"][c]["tour"]
        assert tour == want["tours"][c].tolist(), c
        assert got["items"][c]["cost"] == float(want["costs"][c])
        assert _cycle(tour) <= members[c % count], c
    np.testing.assert_array_equal(got["colony_work"], want["work"])
    assert got["colony_work"].tolist() != ref["colony_work"].tolist()   # the graph changes the search
    # the reference's Hard result keys; rejections per colony and in total equal the oracle's
    assert got["constraint_mode"] == "hard" and got["graph_spec_id"] == 1
    assert got["graph_edges_per_colony"] == [len(members[c % count]) for c in range(4)]
    assert got["reserved_escape_device_bytes"] == 0 and "constraint_mode" not in ref
    np.testing.assert_array_equal(hard.colony_constraint_rejections(4), want["constraint_rejections"])
    assert got["completed_constraint_rejections"] == int(want["constraint_rejections"].sum()) > 0


@pytest.mark.parametrize("mode", ["program", "baseline"])
def test_controlled_hard_runs_match_the_oracle(ext, oracle_mod, mode):
    O = oracle_mod
    n, count, batches = 180, 2, 12
    xys = [(clustered_instance if i % 2 else uniform_instance)(n, seed=47, instance=i) for i in range(count)]
    keys = [400 + i for i in range(count)]
    s = ext.FixedFacoSettings()
    s.ants = 16
    plain = ext.FacoBatchEngine(n, 4, s)
    hard = ext.FacoBatchEngine(n, 4, s, constraint_mode="hard")
    graphs = []
    for k, xy in zip(keys, xys):
        plain.register_problem(k, xy)
        spec, edges = _spec(plain.problem_data(k), [])
        hard.register_graph_problem(k, xy, spec)
        graphs.append(O.csr(sorted(edges), n))
    task_keys = np.array(keys * 2, dtype=np.uint64)
    seeds = np.array([3, 3, 4, 4], dtype=np.uint64)
    program = {"ir_version": 1, "numeric_spec_id": 1, "feature_spec_id": 2, "opcode": [0, 0, 2],
               "operand": [4, 1, 0], "constant_bits": []}          # restarts readily
    baseline = ext.read_baseline({"policy_spec_id": 1, "kind": "rule", "mne_level": 0, "max_mne_level": 3,
                                  "region": 1, "restart_mode": "none", "restart_period": 0, "restart_probability": 0.0,
                                  "stagnation_step": 2, "restart_stagnation": 5, "restart_cooldown": 3})
    announce("hard-control", mode)
    kw = {"programs": [ext.read_program(program)]} if mode == "program" else {"baseline": baseline}
    raw = hard.evaluate_control_raw(task_keys, seeds, s.ants * batches, mode, trace_batches=range(batches), **kw)
    settings = O.default_settings(**{f: getattr(s, f) for f in (
        "ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
        "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")})
    okw = ({"programs": [O.Program.from_buffer_copy(bytes(kw["programs"][0]))]} if mode == "program"
           else {"baseline": O.Baseline.from_buffer_copy(bytes(baseline))})
    want = O.run_controlled(np.stack(xys).astype(np.int32), np.array(keys, dtype=np.uint64), seeds, settings,
                            s.ants * batches, mode, trace=True, graphs=graphs, **okw)
    assert raw["tours"].tolist() == want["tours"].tolist()
    np.testing.assert_array_equal(raw["work"], want["work"])
    np.testing.assert_array_equal(raw["trace_actions"], want["trace_actions"])
    assert [st["restarts"] for st in raw["states"]] == [st["restarts"] for st in want["states"]]
    np.testing.assert_array_equal(raw["rejections"], want["constraint_rejections"])
    assert raw["result"].completed_constraint_rejections == int(want["constraint_rejections"].sum()) > 0


def test_hard_registration_rules(ext):
    n = 120
    xy = uniform_instance(n, seed=43)
    s = ext.FixedFacoSettings()
    plain = ext.FacoBatchEngine(n, 2, s)
    plain.register_problem(1, xy)
    data = plain.problem_data(1)
    spec, _ = _spec(data, [])
    hard = ext.FacoBatchEngine(n, 2, s, constraint_mode="hard")
    with pytest.raises(ValueError):
        hard.register_problem(1, xy)                      # Hard instances come with their graph
    missing = dict(spec, edges=spec["edges"][1:])          # an initial-tour or row edge dropped
    with pytest.raises(ValueError):
        hard.register_graph_problem(1, xy, missing)
    other = dict(spec, common_initial_tour=spec["common_initial_tour"][::-1])
    with pytest.raises(ValueError):
        hard.register_graph_problem(1, xy, other)
    looped = dict(spec, edges=spec["edges"] + [[3, 3]])
    with pytest.raises(ValueError):
        hard.register_graph_problem(1, xy, looped)
    hard.register_graph_problem(1, xy, spec)
    hard.register_graph_problem(1, xy, spec)              # the same registration again
    with pytest.raises(ValueError):
        hard.set_local_search("2opt+oropt")
    with pytest.raises(ValueError):
        ext.FacoBatchEngine(n, 2, s, constraint_mode="sideways")



def test_hard_engines_take_evaluation_counts_only(ext):
    """The reference's Hard engine refuses time budgets, whose cheap tour may
    leave the graph (cuda/batch_engine.cu:539-540)."""
    n = 150
    xy = uniform_instance(n, seed=53)
    s = ext.FixedFacoSettings()
    s.ants = 16
    plain = ext.FacoBatchEngine(n, 2, s)
    plain.register_problem(7, xy)
    spec, _ = _spec(plain.problem_data(7), [])
    hard = ext.FacoBatchEngine(n, 2, s, constraint_mode="hard")
    hard.register_graph_problem(7, xy, spec)
    keys, seeds = np.array([7, 7], dtype=np.uint64), np.array([1, 2], dtype=np.uint64)
    announce("hard-deadline")
    with pytest.raises(ValueError):
        hard.evaluate(keys, seeds, 0.2, 8)
    out = hard.evaluate_faco_evaluations(keys, seeds, s.ants * 3)
    assert out["constraint_mode"] == "hard" and out["completed_constraint_rejections"] == int(
        hard.colony_constraint_rejections(2).sum())


def _independent_spec(O, data, xy, rng, pw, bw, lw):
    """A graph spec with rows of its own (reference prepared_problem.cpp:128-166):
    the graph is the initial tour, a random half of each city's 12 nearest
    edges and a few random chords; primary rows are graph neighbours in a
    random order, LS rows graph neighbours by (distance, city), backup rows
    random other cities; every row right-padded with n."""
    n = xy.shape[0]
    M = O.distance_matrix(xy.astype(np.int64))
    tour = [int(v) for v in data["initial_tour"]]
    edges = set(_cycle(tour))
    for a in range(n):
        near = [b for b in np.argsort(M[a], kind="stable") if b != a][:12]
        for b in near:
            if rng.random() < 0.5:
                edges.add(tuple(sorted((a, int(b)))))
    for _ in range(n // 8):
        a, b = (int(v) for v in rng.choice(n, 2, replace=False))
        edges.add(tuple(sorted((a, b))))
    nbrs = [[] for _ in range(n)]
    for a, b in edges:
        nbrs[a].append(b)
        nbrs[b].append(a)
    primary = np.full((n, pw), n, dtype=np.int64)
    ls = np.full((n, lw), n, dtype=np.int64)
    backup = np.full((n, bw), n, dtype=np.int64)
    for a in range(n):
        chosen = list(rng.permutation(nbrs[a]))[:pw]
        primary[a, :len(chosen)] = chosen
        by_distance = sorted(nbrs[a], key=lambda b: (M[a, b], b))[:lw]
        ls[a, :len(by_distance)] = by_distance
        others = [int(b) for b in rng.permutation(n) if b != a and b not in chosen][:bw - int(rng.integers(0, 3))]
        backup[a, :len(others)] = others
    assert len(edges) <= n * (pw + 1)
    spec = {"graph_spec_id": 1, "common_initial_tour": tour, "edges": [list(e) for e in sorted(edges)],
            "primary": primary.tolist(), "backup": backup.tolist(), "ls": ls.tolist()}
    return spec, edges, (primary, backup, ls)


@pytest.mark.parametrize("n", [200, 1100])
def test_hard_mode_with_independent_rows_matches_the_oracle(ext, oracle_mod, n):
    """General Hard-mode graphs: the spec's primary, backup and LS rows
    replace the kNN rows (split neighbour rows on the device, padding
    included); plain and GP-controlled runs equal the oracle given the same
    rows, every tour inside its graph."""
    O = oracle_mod
    count, batches = 2, 5
    xys = [(clustered_instance if i % 2 else uniform_instance)(n, seed=59, instance=i) for i in range(count)]
    keys = [700 + i for i in range(count)]
    s = ext.FixedFacoSettings()
    s.ants = 32
    s.backup_width = 8
    plain = ext.FacoBatchEngine(n, 4, s)
    hard = ext.FacoBatchEngine(n, 4, s, constraint_mode="hard")
    rng = np.random.default_rng(n + 1)
    graphs, members, rows = [], [], []
    for k, xy in zip(keys, xys):
        plain.register_problem(k, xy)
        spec, edges, triple = _independent_spec(O, plain.problem_data(k), xy, rng, s.primary_width, s.backup_width,
                                                s.ls_width)
        hard.register_graph_problem(k, xy, spec)
        data = hard.problem_data(k)
        assert data["primary"].tolist() == triple[0].tolist() and data["ls"].tolist() == triple[2].tolist()
        graphs.append(O.csr(sorted(edges), n))
        members.append(edges)
        rows.append(triple)
    task_keys = np.array([keys[c % count] for c in range(4)], dtype=np.uint64)
    seeds = np.array([1, 1, 2, 2], dtype=np.uint64)
    settings = O.default_settings(**{f: getattr(s, f) for f in (
        "ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
        "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")})
    announce("hard-rows", n)
    got = hard.evaluate_faco_evaluations(task_keys, seeds, s.ants * batches)
    want = O.run_faco(np.stack(xys).astype(np.int32), keys, seeds, settings, batches, 8, graphs=graphs, rows=rows)
    for c in range(4):
        tour = got["items"][c]["tour"]
        assert tour == want["tours"][c].tolist(), c
        assert got["items"][c]["cost"] == float(want["costs"][c])
        assert _cycle(tour) <= members[c % count], c
    np.testing.assert_array_equal(got["colony_work"], want["work"])
    np.testing.assert_array_equal(hard.colony_constraint_rejections(4), want["constraint_rejections"])
    program = {"ir_version": 1, "numeric_spec_id": 1, "feature_spec_id": 2, "opcode": [0, 0, 2],
               "operand": [4, 1, 0], "constant_bits": []}
    raw = hard.evaluate_control_raw(task_keys, seeds, s.ants * batches * 2, "program",
                                    programs=[ext.read_program(program)])
    wantc = O.run_controlled(np.stack(xys).astype(np.int32), np.array(keys, dtype=np.uint64), seeds, settings,
                             s.ants * batches * 2, "program", graphs=graphs, rows=rows,
                             programs=[O.Program.from_buffer_copy(bytes(ext.read_program(program)))])
    assert raw["tours"].tolist() == wantc["tours"].tolist()
    np.testing.assert_array_equal(raw["work"], wantc["work"])
    assert [st["restarts"] for st in raw["states"]] == [st["restarts"] for st in wantc["states"]]


def test_independent_row_rules(ext, oracle_mod):
    """The reference's row checks (prepared_problem.cpp:139-162): fixed
    widths, cities before padding, no repeats, primary and LS members in the
    graph, LS rows ordered by distance, backup rows disjoint from primary
    rows; the same key again with other rows is refused. Rows hold EUC_2D
    distances, so two LS cities of one rounded distance may come in either
    order (an exact-distance order stays valid after rounding)."""
    n = 90
    xy = uniform_instance(n, seed=61)
    s = ext.FixedFacoSettings()
    s.backup_width = 4
    plain = ext.FacoBatchEngine(n, 2, s)
    plain.register_problem(1, xy)
    spec, _, (primary, backup, ls) = _independent_spec(oracle_mod, plain.problem_data(1), xy,
                                                       np.random.default_rng(3), s.primary_width, s.backup_width,
                                                       s.ls_width)
    hard = ext.FacoBatchEngine(n, 2, s, constraint_mode="hard")

    def variant(name, rows):
        return dict(spec, **{name: rows.tolist()})
    def euc(i, j):
        return int(np.floor(np.hypot(*(xy[i] - xy[j])) + 0.5))
    a = next(i for i in range(n) if ls[i, 1] < n and euc(i, ls[i, 0]) < euc(i, ls[i, 1]))
    swapped = ls.copy()
    swapped[a, [0, 1]] = swapped[a, [1, 0]]
    gap = primary.copy()
    gap[a, 0] = n                                      # padding before a city
    stranger = primary.copy()
    outside = [b for b in range(n) if b != a and [min(a, b), max(a, b)] not in spec["edges"]][0]
    stranger[a, 0] = outside                           # a primary city outside the graph
    echo = backup.copy()
    echo[a, 0] = primary[a, 0]                         # a backup city repeating a primary one
    for bad in (variant("ls", swapped), variant("primary", gap), variant("primary", stranger),
                variant("backup", echo)):
        with pytest.raises(ValueError):
            hard.register_graph_problem(1, xy, bad)
    hard.register_graph_problem(1, xy, spec)
    with pytest.raises(ValueError):
        hard.register_graph_problem(1, xy, variant("backup", np.roll(backup, 1, axis=1)))


def test_matched_prior_graphs_run_in_hard_mode_like_the_oracle(ext, oracle_mod):
    """Graphs matched from two candidate priors (candidate_prior.match_graphs,
    reference graph_matching.py:37-169) register in a Hard engine as they
    come (engine_graph_spec) and run like the oracle given the same graph and
    rows: LS rows ordered by exact distance, primary rows by prior rank,
    native backup tails and uniform backup slots, padding where a node has
    fewer graph neighbours than the widths."""
    from types import SimpleNamespace

    from gplsaco_amd.candidate_prior import GraphSettings, engine_graph_spec, match_graphs
    from tests.test_candidate_prior import matched_priors
    O = oracle_mod
    n, batches = 200, 4
    xy = clustered_instance(n, seed=67, instance=0)
    s = ext.FixedFacoSettings()
    s.ants = 32
    s.backup_width = 8
    plain = ext.FacoBatchEngine(n, 2, s)
    plain.register_problem(1, xy)
    tour = tuple(int(v) for v in plain.problem_data(1)["initial_tour"])
    problem = SimpleNamespace(instance_id="clustered-200", numeric_id=1, coordinates=tuple(map(tuple, xy.tolist())))
    graphs = match_graphs(problem, tour, matched_priors(problem, 24, 11),
                          GraphSettings(primary_width=s.primary_width, backup_width=8, ls_width=s.ls_width,
                                        uniform_backup_slots=4))
    settings = O.default_settings(**{f: getattr(s, f) for f in (
        "ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
        "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")})
    for kind, graph in graphs.items():
        hard = ext.FacoBatchEngine(n, 2, s, constraint_mode="hard")
        hard.register_graph_problem(1, xy, engine_graph_spec(graph))
        rows = tuple(np.asarray(graph[name], dtype=np.int64) for name in ("primary", "backup", "ls"))
        edges = {tuple(e) for e in graph["edges"]}
        seeds = np.array([3, 4], dtype=np.uint64)
        announce("hard-matched", kind)
        got = hard.evaluate_faco_evaluations(np.array([1, 1], dtype=np.uint64), seeds, s.ants * batches)
        want = O.run_faco(xy[None].astype(np.int32), [1], seeds, settings, batches, 8,
                          graphs=[O.csr(sorted(edges), n)], rows=[rows])
        for c in range(2):
            assert got["items"][c]["tour"] == want["tours"][c].tolist(), (kind, c)
            assert _cycle(got["items"][c]["tour"]) <= edges
        np.testing.assert_array_equal(got["colony_work"], want["work"])
        np.testing.assert_array_equal(hard.colony_constraint_rejections(2), want["constraint_rejections"])


def test_hard_rejections_of_a_node_cut_short_by_the_limit(ext, oracle_mod):
    """A node the LS evaluation limit cuts short still counts the rejected
    moves it evaluated before the limit, as the reference's sequential scan
    does (cuda/faco_device.cuh:316-336): small limits stop most ants inside
    a node, and device and oracle agree on tours, work and rejections."""
    O = oracle_mod
    n, batches = 200, 4
    xy = clustered_instance(n, seed=71, instance=0)
    s = ext.FixedFacoSettings()
    s.ants = 16
    plain = ext.FacoBatchEngine(n, 2, s)
    plain.register_problem(1, xy)
    spec, edges = _spec(plain.problem_data(1), [])
    graph = O.csr(sorted(edges), n)
    seeds = np.array([8, 9], dtype=np.uint64)
    for limit in (7, 37, 101):
        s.ls_evaluation_limit = limit
        hard = ext.FacoBatchEngine(n, 2, s, constraint_mode="hard")
        hard.register_graph_problem(1, xy, spec)
        settings = O.default_settings(**{f: getattr(s, f) for f in (
            "ants", "primary_width", "backup_width", "ls_width", "beta", "retention", "p_best",
            "epoch_source_probability", "ls_evaluation_limit", "initial_ls_evaluation_limit")})
        announce("hard-cut", limit)
        got = hard.evaluate_faco_evaluations(np.array([1, 1], dtype=np.uint64), seeds, s.ants * batches)
        want = O.run_faco(xy[None].astype(np.int32), [1], seeds, settings, batches, 8, graphs=[graph])
        for c in range(2):
            assert got["items"][c]["tour"] == want["tours"][c].tolist(), (limit, c)
        np.testing.assert_array_equal(got["colony_work"], want["work"])
        np.testing.assert_array_equal(hard.colony_constraint_rejections(2), want["constraint_rejections"])
        assert want["work"][:, 1].tolist() == [limit * s.ants * batches] * 2 or limit > 100
