"""Candidate priors from an external tool (LKH's ALPHA or POPMUSIC
candidate sets) turned into Hard-mode graph specs: the counterpart of the
reference's python/gp_faco/candidate_prior.py (parse_candidates) and
python/gp_faco/graph_matching.py (match_graphs, engine_graph_spec).

The engine never runs the external tool; it takes the tool's candidate
export (format version 1: per node its id, pi, MST parent, count and
(destination, alpha, transformed cost, quantized distance) edges), checks
it against the instance's own coordinates, and matches the two priors of an
instance to graphs of the same edge count and the same actual slots per
node, so that runs under either prior differ only in which edges they may
use. The result feeds FacoBatchEngine.register_graph_problem unchanged.
"""
from __future__ import annotations

import math
import random
from dataclasses import asdict, dataclass

KINDS = ("ALPHA", "POPMUSIC")


@dataclass(frozen=True)
class PriorSettings:
    """The external tool's candidate settings (reference
    candidate_prior.py:20-41); positive int32 parameters."""
    kind: str = "ALPHA"
    maximum_candidates: int = 80
    seed: int = 73001
    distance_scale: int = 1000000
    popmusic_solutions: int = 50
    popmusic_sample_size: int = 10
    popmusic_max_neighbors: int = 5
    popmusic_trials: int = 1

    def __post_init__(self):
        if self.kind not in KINDS:
            raise ValueError("the prior is ALPHA or POPMUSIC")
        for name, value in asdict(self).items():
            if name != "kind" and (type(value) is not int or not 1 <= value <= 0x7FFFFFFF):
                raise ValueError(f"{name} must be a positive int32")
        if self.maximum_candidates > 9999:
            raise ValueError("at most 9999 candidates per node")


def _coordinates(problem):
    return [tuple(map(float, xy)) for xy in problem.coordinates]


def _euclidean(xy, i, j):
    dx, dy = xy[i][0] - xy[j][0], xy[i][1] - xy[j][1]
    return math.sqrt(dx * dx + dy * dy)


def parse_candidates(problem, settings: PriorSettings, raw: dict) -> dict:
    """Check a candidate export against the instance (reference
    candidate_prior.py:44-135) and return the prior: per node the tool's
    candidates in its own priority order (alpha, transformed cost, quantized
    and FP64 distance) and the same members by FP64 distance. Every id,
    count and cost relation is checked: 1-based nodes in order, no self
    loops or repeats, quantized = round(scale * distance), transformed =
    quantized + pi(i) + pi(j)."""
    n = len(problem.coordinates)
    fields = {"candidate_export_version", "dimension", "kind", "scale", "precision", "seed", "maximum_candidates",
              "native_preparation_cpu_seconds", "nodes"}
    if not isinstance(raw, dict) or set(raw) != fields:
        raise ValueError("unknown fields in the candidate export")
    expected = {"candidate_export_version": 1, "dimension": n, "kind": settings.kind, "scale": settings.distance_scale,
                "precision": 1, "seed": settings.seed, "maximum_candidates": settings.maximum_candidates}
    for key, value in expected.items():
        if type(raw[key]) is not type(value) or raw[key] != value:
            raise ValueError(f"the candidate export does not match the request: {key}")
    seconds = raw["native_preparation_cpu_seconds"]
    if type(seconds) not in (int, float) or not math.isfinite(seconds) or seconds < 0:
        raise ValueError("invalid preparation time in the candidate export")
    nodes = raw["nodes"]
    if type(nodes) is not list or len(nodes) != n:
        raise ValueError("the candidate export does not cover every node")
    for i, node in enumerate(nodes):
        if (not isinstance(node, dict) or set(node) != {"id", "pi", "dad", "count", "edges"}
                or any(type(node[k]) is not int for k in ("id", "pi", "dad", "count"))
                or node["id"] != i + 1 or not 0 <= node["dad"] <= n):
            raise ValueError("node id, pi or MST parent invalid")
        if not 0 <= node["count"] < n or type(node["edges"]) is not list or len(node["edges"]) != node["count"]:
            raise ValueError("node candidate count does not match its edges")
    xy = _coordinates(problem)
    rows, distance_rows, undirected = [], [], set()
    worst = 0.0
    for i, node in enumerate(nodes):
        row, members = [], set()
        for edge in node["edges"]:
            if type(edge) is not list or len(edge) != 4 or any(type(v) is not int for v in edge):
                raise ValueError("a candidate edge is four integers")
            destination, alpha, transformed, quantized = edge
            j = destination - 1
            if not 0 <= j < n or j == i or j in members:
                raise ValueError("a candidate is unknown, the node itself or repeated")
            members.add(j)
            distance = _euclidean(xy, i, j)
            if quantized != math.floor(settings.distance_scale * distance + 0.5):
                raise ValueError("a quantized distance does not match the coordinates")
            if transformed != quantized + node["pi"] + nodes[j]["pi"]:
                raise ValueError("a transformed cost does not match distance and pi")
            worst = max(worst, abs(quantized / settings.distance_scale - distance))
            row.append({"to": j, "alpha": alpha, "transformed_cost": transformed, "quantized_distance": quantized,
                        "distance": distance})
            undirected.add((min(i, j), max(i, j)))
        rows.append(row)
        distance_rows.append([e["to"] for e in sorted(row, key=lambda e: (e["distance"], e["to"]))])
    degrees = [len(r) for r in rows]
    return {"prior_spec_id": 1, "dimension": n, "instance_id": problem.instance_id, "settings": asdict(settings),
            "rows": rows, "distance_rows": distance_rows, "node_pi": [v["pi"] for v in nodes],
            "directed_slots": sum(degrees), "undirected_edges": len(undirected), "degree_min": min(degrees),
            "degree_max": max(degrees), "max_distance_quantization_error": worst,
            "distance_role": "the tool's integer distances select members; FP64 source distances order LS rows",
            "initial_tour_exported": False}


@dataclass(frozen=True)
class GraphSettings:
    """Row widths of the matched graphs and the uniform backup slots
    (reference graph_matching.py:20-34)."""
    primary_width: int = 16
    backup_width: int = 64
    ls_width: int = 20
    uniform_backup_slots: int = 16
    preparation_seed: int = 80471

    def __post_init__(self):
        if (any(type(v) is not int or v < 0 for v in asdict(self).values()) or self.primary_width < 2
                or self.ls_width < 1 or self.uniform_backup_slots > self.backup_width):
            raise ValueError("invalid row widths, uniform slots or preparation seed")


def engine_graph_spec(graph: dict) -> dict:
    """The six fields register_graph_problem takes (graph_matching.py:12-17)."""
    return {k: graph[k] for k in ("graph_spec_id", "common_initial_tour", "edges", "primary", "backup", "ls")}


def match_graphs(problem, initial_tour, priors: dict, settings: GraphSettings | None = None) -> dict:
    """Both priors' graphs of one instance, matched (reference
    graph_matching.py:37-169):

    - edges: the initial tour's plus, per prior, its extra candidate edges
      (within the first primary_width of a node's row, outside the tour)
      ranked by (better rank from either end, worse rank, a, b), both priors
      cut to the smaller count;
    - per node the same actual slots under both priors: primary = the
      node's graph neighbours by prior rank (ties by city), LS = the
      neighbours by FP64 distance, both as many as the smaller degree allows;
      backup = the prior's row tail past primary_width (not in the primary
      row), as many as both priors have, then uniform_backup_slots cities
      drawn from a Random(preparation seed << 96 | numeric id << 32 | node)
      stream; every row right-padded with n."""
    settings = settings or GraphSettings()
    if type(settings) is not GraphSettings:
        raise TypeError("graph settings must be GraphSettings")
    n = len(problem.coordinates)
    tour = [int(v) for v in initial_tour]
    if len(tour) != n or sorted(tour) != list(range(n)):
        raise ValueError("the initial tour must visit every city once")
    if set(priors) != set(KINDS):
        raise ValueError("matching needs both priors")
    width = min(settings.primary_width, n - 1)
    backup_width = min(settings.backup_width, n - 1 - width)
    ls_width = min(settings.ls_width, n - 1)
    tour_edges = {tuple(sorted((tour[i - 1], tour[i]))) for i in range(n)}
    ranks, extras = {}, {}
    for kind, prior in priors.items():
        if (type(prior.get("prior_spec_id")) is not int or prior["prior_spec_id"] != 1
                or type(prior.get("dimension")) is not int or prior["dimension"] != n
                or prior.get("instance_id", problem.instance_id) != problem.instance_id
                or prior["settings"]["kind"] != kind or len(prior["rows"]) != n):
            raise ValueError("a prior does not belong to this instance and method")
        ranks[kind], wanted = [], set()
        for i, row in enumerate(prior["rows"]):
            ids = [e["to"] for e in row]
            if any(type(j) is not int or not 0 <= j < n or j == i for j in ids) or len(ids) != len(set(ids)):
                raise ValueError("a prior row holds an unknown city, the node itself or a repeat")
            ranks[kind].append({j: r for r, j in enumerate(ids)})
            wanted.update(tuple(sorted((i, j))) for j in ids[:width])
        extras[kind] = wanted - tour_edges
    budget = min(len(v) for v in extras.values())
    edges_of, neighbours_of = {}, {}
    for kind in KINDS:
        rank = ranks[kind]

        def priority(edge, rank=rank):
            a, b = edge
            ra, rb = rank[a].get(b, n), rank[b].get(a, n)
            return min(ra, rb), max(ra, rb), a, b
        edges = tour_edges | set(sorted(extras[kind], key=priority)[:budget])
        neighbours = [[] for _ in range(n)]
        for a, b in sorted(edges):
            neighbours[a].append(b)
            neighbours[b].append(a)
        edges_of[kind], neighbours_of[kind] = edges, neighbours
    primary_slots = [min(width, *(len(v[i]) for v in neighbours_of.values())) for i in range(n)]
    ls_slots = [min(ls_width, *(len(v[i]) for v in neighbours_of.values())) for i in range(n)]
    main = {kind: [sorted(row, key=lambda j, i=i: (ranks[kind][i].get(j, n), j))[:primary_slots[i]]
                   for i, row in enumerate(neighbours_of[kind])] for kind in KINDS}
    tails = {kind: [[e["to"] for e in priors[kind]["rows"][i][width:] if e["to"] != i and e["to"] not in main[kind][i]]
                    for i in range(n)] for kind in KINDS}
    uniform_slots = min(settings.uniform_backup_slots, backup_width)
    tail_slots = [min(backup_width - uniform_slots, *(len(t[i]) for t in tails.values())) for i in range(n)]
    xy = _coordinates(problem)
    out = {}
    for kind in KINDS:
        primary, backup, ls = [], [], []
        for i, row in enumerate(neighbours_of[kind]):
            chosen = main[kind][i]
            primary.append(chosen + [n] * (width - len(chosen)))
            by_distance = sorted(row, key=lambda j, i=i: (_euclidean(xy, i, j), j))[:ls_slots[i]]
            ls.append(by_distance + [n] * (ls_width - len(by_distance)))
            tail = tails[kind][i][:tail_slots[i]]
            excluded = {i, *chosen, *tail}
            pool = [j for j in range(n) if j not in excluded]
            stream = random.Random((settings.preparation_seed << 96) | (problem.numeric_id << 32) | i)
            values = tail + stream.sample(pool, uniform_slots)
            backup.append(values + [n] * (backup_width - len(values)))
        graph = {"graph_spec_id": 1, "matching_spec_id": 2, "dimension": n, "instance_id": problem.instance_id,
                 "common_initial_tour": tour, "edges": [list(e) for e in sorted(edges_of[kind])],
                 "primary": primary, "backup": backup, "ls": ls, "settings": asdict(settings), "prior_kind": kind,
                 "prior_id": f"{problem.instance_id}-{kind}", "extra_edge_budget": budget,
                 "discarded_extra_edges": len(extras[kind]) - budget,
                 "degrees": [len(v) for v in neighbours_of[kind]],
                 "matched_actual_slots": {"primary": list(primary_slots), "ls": list(ls_slots),
                                          "native_backup": list(tail_slots), "uniform_backup": [uniform_slots] * n},
                 "matching_policy": "paired-node minimum actual slots; full E0 edges retained",
                 "preparation_random_stream": "Python Random/MT19937; explicit preparation seed, instance ordinal, node",
                 "graph_id": f"{problem.instance_id}-{kind}"}
        out[kind] = graph
    return out
