"""Candidate priors and matched graphs (gplsaco_amd/candidate_prior.py)
against the reference's own fixtures and expectations:
tests/python/test_candidate_prior.py (the triangle export, its FP64 and
priority views and nine malformed variants) and
tests/python/test_graph_matching.py (the seven-city graph with its exact
edge budget and rows, identity checks, and the per-node slot matching of two
priors of unequal degrees). A plausible mistake fails one of them: LS rows
by prior rank instead of distance (ls[0] == [1, 2, 6]), extras ranked by the
worse end (the kept edge (0, 2)), a budget from the larger prior (edge
count), tails without the primary exclusion (disjoint rows), slots matched
per prior instead of per node (equal matched slots).
"""
import copy
import math
import random
from types import SimpleNamespace

import numpy as np
import pytest

from gplsaco_amd.candidate_prior import (GraphSettings, PriorSettings, engine_graph_spec, match_graphs,
                                         parse_candidates)


def triangle():
    problem = SimpleNamespace(instance_id="triangle", numeric_id=0,
                              coordinates=((0.0, 0.0), (0.3, 0.4), (0.0, 1.0)))
    settings = PriorSettings(maximum_candidates=2, distance_scale=10, seed=11)
    raw = {"candidate_export_version": 1, "dimension": 3, "kind": "ALPHA", "scale": 10, "precision": 1, "seed": 11,
           "maximum_candidates": 2, "native_preparation_cpu_seconds": 0.1,
           "nodes": [{"id": 1, "pi": -4, "dad": 0, "count": 2, "edges": [[3, 0, 9, 10], [2, 1, 3, 5]]},
                     {"id": 2, "pi": 2, "dad": 1, "count": 2, "edges": [[1, 0, 3, 5], [3, 0, 12, 7]]},
                     {"id": 3, "pi": 3, "dad": 2, "count": 2, "edges": [[1, 0, 9, 10], [2, 0, 12, 7]]}]}
    return problem, settings, raw


def test_prior_keeps_the_tools_order_and_an_fp64_distance_view():
    problem, settings, raw = triangle()
    prior = parse_candidates(problem, settings, raw)
    assert [e["to"] for e in prior["rows"][0]] == [2, 1]
    assert prior["distance_rows"][0] == [1, 2]
    assert prior["directed_slots"] == 6 and prior["undirected_edges"] == 3
    assert prior["rows"][1][1]["distance"] == pytest.approx(math.sqrt(0.45))
    assert prior["rows"][1][1]["quantized_distance"] == 7
    assert prior["node_pi"] == [-4, 2, 3]
    assert prior["degree_min"] == prior["degree_max"] == 2
    assert prior["max_distance_quantization_error"] == pytest.approx(abs(0.7 - math.sqrt(0.45)))
    assert prior["initial_tour_exported"] is False


@pytest.mark.parametrize("mutation", ["identity", "dimension", "node_id", "duplicate", "quantized", "transformed",
                                      "count", "float_edge", "extra", "self_loop", "bad_time"])
def test_malformed_candidate_exports_are_refused(mutation):
    problem, settings, original = triangle()
    raw = copy.deepcopy(original)
    if mutation == "identity":
        raw["seed"] = 12
    elif mutation == "dimension":
        raw["nodes"].pop()
    elif mutation == "node_id":
        raw["nodes"][1]["id"] = 3
    elif mutation == "duplicate":
        raw["nodes"][0]["edges"][1] = raw["nodes"][0]["edges"][0]
    elif mutation == "quantized":
        raw["nodes"][0]["edges"][0][3] = 11
    elif mutation == "transformed":
        raw["nodes"][0]["edges"][0][2] = 8
    elif mutation == "count":
        raw["nodes"][0]["count"] = 1
    elif mutation == "float_edge":
        raw["nodes"][0]["edges"][0][0] = 3.0
    elif mutation == "self_loop":
        raw["nodes"][0]["edges"][0] = [1, 0, -8, 0]
    elif mutation == "bad_time":
        raw["native_preparation_cpu_seconds"] = float("nan")
    else:
        raw["optimal_tour"] = [1, 2, 3]
    with pytest.raises(ValueError):
        parse_candidates(problem, settings, raw)


def test_prior_settings_bounds():
    with pytest.raises(ValueError):
        PriorSettings(kind="NEAREST")
    with pytest.raises(ValueError):
        PriorSettings(maximum_candidates=10000)
    with pytest.raises(ValueError):
        PriorSettings(seed=0)
    with pytest.raises(ValueError):
        GraphSettings(primary_width=1)
    with pytest.raises(ValueError):
        GraphSettings(backup_width=4, uniform_backup_slots=5)


def seven():
    problem = SimpleNamespace(instance_id="seven", numeric_id=1,
                              coordinates=((0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (2.0, 0.5), (3.0, 1.0), (2.0, 2.0),
                                           (0.0, 2.0)))
    priors = {}
    for kind in ("ALPHA", "POPMUSIC"):
        rows = []
        for i in range(7):
            if kind == "ALPHA":
                ids = [(i + 2) % 7, (i + 1) % 7]
                ids += [j for j in range(7) if j != i and j not in ids]
            else:
                ids = {0: [2, 1], 2: [0, 1]}.get(i, [(i - 1) % 7, (i + 1) % 7])
            rows.append([{"to": j} for j in ids])
        priors[kind] = {"prior_spec_id": 1, "dimension": 7, "instance_id": problem.instance_id,
                        "settings": {"kind": kind}, "rows": rows}
    return problem, priors


def test_seven_city_graph_budget_priority_and_fixed_slots():
    problem, priors = seven()
    settings = GraphSettings(primary_width=2, backup_width=2, ls_width=3, uniform_backup_slots=1)
    result = match_graphs(problem, tuple(range(7)), priors, settings)
    expected = {tuple(sorted((i, (i + 1) % 7))) for i in range(7)} | {(0, 2)}
    for kind, graph in result.items():
        assert set(map(tuple, graph["edges"])) == expected
        assert graph["extra_edge_budget"] == 1
        assert graph["discarded_extra_edges"] == (6 if kind == "ALPHA" else 0)
        assert graph["primary"][0] == [2, 1]
        assert graph["ls"][0] == [1, 2, 6]       # edge 0-6 is not in the width-2 primary row but stays in the graph
        assert graph["degrees"][0] == 3
        assert sum(graph["degrees"]) == 2 * len(expected)
        for i in range(7):
            assert len(graph["primary"][i]) == len(graph["backup"][i]) == 2
            assert len(graph["ls"][i]) == 3
            assert i not in graph["primary"][i] and i not in graph["backup"][i]
            assert not (set(graph["primary"][i]) & set(graph["backup"][i]) - {7})
        assert set(engine_graph_spec(graph)) == {"graph_spec_id", "common_initial_tour", "edges", "primary",
                                                  "backup", "ls"}
    assert match_graphs(problem, tuple(range(7)), priors, settings) == result
    changed = match_graphs(problem, tuple(range(7)), priors,
                           GraphSettings(primary_width=2, backup_width=2, ls_width=3, uniform_backup_slots=1,
                                         preparation_seed=811))
    assert all(changed[k]["edges"] == result[k]["edges"] for k in result)
    assert all(changed[k]["settings"] != result[k]["settings"] for k in result)


def test_matching_refuses_foreign_priors_repeats_and_bad_tours():
    problem, priors = seven()
    for kind in ("ALPHA", "POPMUSIC"):
        wrong = copy.deepcopy(priors)
        wrong[kind]["instance_id"] = "foreign"
        with pytest.raises(ValueError):
            match_graphs(problem, tuple(range(7)), wrong)
        wrong = copy.deepcopy(priors)
        wrong[kind]["rows"][0].append(wrong[kind]["rows"][0][0])
        with pytest.raises(ValueError):
            match_graphs(problem, tuple(range(7)), wrong)
    with pytest.raises(ValueError):
        match_graphs(problem, (0, 0, 2, 3, 4, 5, 6), priors)
    with pytest.raises(ValueError):
        match_graphs(problem, tuple(range(7)), {"ALPHA": priors["ALPHA"]})
    with pytest.raises(TypeError):
        match_graphs(problem, tuple(range(7)), priors, {"primary_width": 2})


def test_actual_slots_match_per_node_with_unequal_degrees_and_prior_tails():
    n = 15
    problem = SimpleNamespace(instance_id="unequal", numeric_id=0,
                              coordinates=tuple((float(i * i), float(i % 4)) for i in range(n)))
    rng, priors = random.Random(919), {}
    for kind in ("ALPHA", "POPMUSIC"):
        rows = []
        for i in range(n):
            values = rng.sample([j for j in range(n) if j != i], n - 1)
            if kind == "POPMUSIC":
                values = values[: 2 + i % 8]
            rows.append([{"to": j} for j in values])
        priors[kind] = {"prior_spec_id": 1, "dimension": n, "instance_id": problem.instance_id,
                        "settings": {"kind": kind}, "rows": rows}
    settings = GraphSettings(primary_width=4, backup_width=6, ls_width=3, uniform_backup_slots=2)
    result = match_graphs(problem, tuple(range(n)), priors, settings)
    a, b = result["ALPHA"], result["POPMUSIC"]
    assert len(a["edges"]) == len(b["edges"])
    assert a["degrees"] != b["degrees"]
    assert a["matched_actual_slots"] == b["matched_actual_slots"]
    assert any(v == 0 for v in a["matched_actual_slots"]["native_backup"])
    assert any(v > 0 for v in a["matched_actual_slots"]["native_backup"])
    for i in range(n):
        primary_count = min(4, a["degrees"][i], b["degrees"][i])
        ls_count = min(3, a["degrees"][i], b["degrees"][i])
        tails = []
        for kind, graph in result.items():
            main = [j for j in graph["primary"][i] if j != n]
            tails.append([e["to"] for e in priors[kind]["rows"][i][4:] if e["to"] not in {i, *main}])
        tail_count = min(4, *(len(v) for v in tails))
        for kind, graph in result.items():
            assert graph["matching_spec_id"] == 2
            assert sum(j != n for j in graph["primary"][i]) == primary_count
            assert sum(j != n for j in graph["ls"][i]) == ls_count
            assert sum(j != n for j in graph["backup"][i]) == tail_count + 2
            assert len(set(graph["backup"][i]) - {n}) == tail_count + 2
            assert not (set(graph["primary"][i]) & set(graph["backup"][i]) - {n})
            # the native tail leads the backup row, in the prior's order
            assert graph["backup"][i][:tail_count] == tails[0 if kind == "ALPHA" else 1][:tail_count]
            neighbours = [y if x == i else x for x, y in graph["edges"] if i in (x, y)]
            assert graph["ls"][i][:ls_count] == sorted(
                neighbours, key=lambda j: (math.dist(problem.coordinates[i], problem.coordinates[j]), j))[:ls_count]
            assert [j for j in graph["primary"][i] if j != n] == sorted(
                neighbours, key=lambda j: (next((r for r, e in enumerate(priors[kind]["rows"][i]) if e["to"] == j), n),
                                           j))[:primary_count]
    assert match_graphs(problem, tuple(range(n)), priors, settings) == result


def synthetic_export(xy, kind, settings, rng, width):
    """A candidate export of the tool's format for integer coordinates:
    ALPHA rows are the nearest cities by distance, POPMUSIC rows a shuffled
    sample of the 2 width nearest; pi and alpha random integers."""
    n = xy.shape[0]
    d = np.hypot(*(xy[:, None, :] - xy[None, :, :]).transpose(2, 0, 1))
    pi = rng.integers(-50, 50, n)
    nodes = []
    for i in range(n):
        near = [int(j) for j in np.argsort(d[i], kind="stable") if j != i]
        ids = near[:width] if kind == "ALPHA" else list(rng.permutation(near[:2 * width])[:width])
        edges = []
        for j in ids:
            q = math.floor(settings.distance_scale * math.sqrt((xy[i, 0] - xy[j, 0]) ** 2 + (xy[i, 1] - xy[j, 1]) ** 2)
                           + 0.5)
            edges.append([int(j) + 1, int(rng.integers(0, 9)), int(q + pi[i] + pi[j]), int(q)])
        nodes.append({"id": i + 1, "pi": int(pi[i]), "dad": 0 if i == 0 else i, "count": len(edges), "edges": edges})
    return {"candidate_export_version": 1, "dimension": n, "kind": kind, "scale": settings.distance_scale,
            "precision": 1, "seed": settings.seed, "maximum_candidates": settings.maximum_candidates,
            "native_preparation_cpu_seconds": 0.0, "nodes": nodes}


def matched_priors(problem, width, seed):
    """Both priors of an instance, parsed from synthetic exports."""
    xy = np.asarray(problem.coordinates, dtype=np.float64)
    rng = np.random.default_rng(seed)
    priors = {}
    for kind in ("ALPHA", "POPMUSIC"):
        settings = PriorSettings(kind=kind, maximum_candidates=width)
        priors[kind] = parse_candidates(problem, settings, synthetic_export(xy, kind, settings, rng, width))
    return priors


def test_synthetic_exports_parse_and_match_into_valid_specs():
    """Parsed priors of a 60-city instance match into specs that satisfy the
    rules a Hard engine checks at registration: the graph holds the initial
    tour and at most n (primary_width + 1) edges, primary and LS rows are
    graph neighbours, LS rows by FP64 distance, backup rows disjoint from
    primary rows, every row right-padded with n."""
    rng = np.random.default_rng(5)
    n = 60
    problem = SimpleNamespace(instance_id="synthetic-60", numeric_id=9,
                              coordinates=tuple(map(tuple, rng.integers(0, 1000, (n, 2)).astype(float).tolist())))
    priors = matched_priors(problem, 12, 3)
    tour = tuple(int(v) for v in rng.permutation(n))
    settings = GraphSettings(primary_width=8, backup_width=10, ls_width=6, uniform_backup_slots=4)
    graphs = match_graphs(problem, tour, priors, settings)
    for graph in graphs.values():
        edges = {tuple(e) for e in graph["edges"]}
        assert {tuple(sorted((tour[i - 1], tour[i]))) for i in range(n)} <= edges
        assert len(edges) <= n * (8 + 1)
        for i in range(n):
            neighbours = {b if a == i else a for a, b in edges if i in (a, b)}
            for name, width in (("primary", 8), ("backup", 10), ("ls", 6)):
                row = graph[name][i]
                assert len(row) == width
                cities = [j for j in row if j != n]
                assert row == cities + [n] * (width - len(cities))
                assert len(set(cities)) == len(cities) and i not in cities
            assert set(graph["primary"][i]) - {n} <= neighbours and set(graph["ls"][i]) - {n} <= neighbours
            ls = [j for j in graph["ls"][i] if j != n]
            assert ls == sorted(ls, key=lambda j: (math.dist(problem.coordinates[i], problem.coordinates[j]), j))
            assert not (set(graph["primary"][i]) & set(graph["backup"][i]) - {n})
