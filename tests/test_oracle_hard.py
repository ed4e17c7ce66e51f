"""The oracle's Hard-mode move predicates (reference
cpp/include/gp_faco/edge_constraints.hpp:47-90), pinned the way the
reference's tests/cpp/test_hard_edges.cpp pins them: against the complete
edge sets of the tour before and after the actual move, never against the
three-edge / two-edge formulas themselves.

For every tour of n = 3..6 cities (city 0 first, both directions), every
relocation (node moved to follow target) and every 2-opt move of both kinds
is carried out on a Python list; the move is allowed in a graph exactly
when the graph holds every edge of the tour after it. The graphs: the
tour's own edges, those plus every new edge of the move, and the latter
less each new edge in turn. Plus the reference's regression cases (a chord
that covers only one of the new edges) and the undirected CSR rules
(symmetry, no truncation of degrees, out-of-range cities never members).

The oracle's Hard-mode runs (cuda/faco_choices.cuh:46-53: a city is
available only when its relocation stays in the graph; faco_device.cuh:383-394:
a 2-opt move leaving the graph never wins): selection exhausts instead of
leaving the graph (the reference's 6-city regression), a run on the
complete graph equals plain FACO bit for bit, and every tour of a run on the
graph of primary candidates plus the initial tour stays inside that graph.
"""
import itertools

import numpy as np

from oracle import oracle as O


def cycle_edges(tour):
    return {tuple(sorted((tour[i], tour[(i + 1) % len(tour)]))) for i in range(len(tour))}


def relocated(tour, target, node):
    order = [v for v in tour if v != node]
    at = order.index(target)
    return order[:at + 1] + [node] + order[at + 1:]


def two_opt_after(tour, a, b, kind):
    """Kind 0 removes (a, succ a), (b, succ b); kind 1 (pred a, a), (pred b, b).
    The segment between the removed edges is reversed as a list."""
    n = len(tour)
    if kind == 0:
        start = (tour.index(a) + 1) % n              # succ a ... b, then succ b ... a
        order = tour[start:] + tour[:start]
        cut = order.index(b) + 1
    else:
        start = tour.index(a)                        # a ... pred b, then b ... pred a
        order = tour[start:] + tour[:start]
        cut = order.index(b)
    return order[:cut][::-1] + order[cut:]


def allowed_by_edges(after, members):
    return cycle_edges(after) <= members


def check_against_graphs(before, after, predicate):
    old, new = cycle_edges(before), cycle_edges(after)
    graphs = [old, old | new] + [(old | new) - {e} for e in new - old]
    rejected = 0
    for members in graphs:
        graph = O.csr(sorted(members), len(before))
        expected = allowed_by_edges(after, members)
        assert predicate(graph) == expected, (before, after, sorted(members))
        rejected += not expected
    return rejected


def test_every_move_of_small_tours_matches_its_edge_sets():
    moves = rejected = 0
    for n in range(3, 7):
        for rest in itertools.permutations(range(1, n)):
            tour = [0, *rest]
            for a in range(n):
                for b in range(n):
                    if a == b:
                        continue
                    rejected += check_against_graphs(tour, relocated(tour, a, b),
                                                     lambda g: O.hard_relocation_allowed(tour, a, b, g))
                    for kind in (0, 1):
                        i, j = tour.index(a), tour.index(b)
                        an = tour[(i + 1) % n] if kind == 0 else tour[i - 1]
                        bn = tour[(j + 1) % n] if kind == 0 else tour[j - 1]
                        rejected += check_against_graphs(tour, two_opt_after(tour, a, b, kind),
                                                         lambda g: O.hard_two_opt_allowed(a, an, b, bn, g, n))
                    moves += 3
    assert moves > 3000 and rejected > 0


def test_reference_regressions():
    order = [0, 1, 2, 3, 4, 5]
    graph = O.csr(sorted(cycle_edges(order) | {(0, 3)}), 6)
    # the chord (0, 3) covers one new edge of the move, not the others
    assert not O.hard_relocation_allowed(order, 0, 3, graph)
    assert not O.hard_two_opt_allowed(0, 1, 3, 4, graph, 6)
    # identity moves stay allowed in the tour's own graph
    ring = O.csr(sorted(cycle_edges(order)), 6)
    assert O.hard_relocation_allowed(order, 0, 1, ring) and O.hard_two_opt_allowed(0, 1, 1, 2, ring, 6)


def test_undirected_graph_rules():
    # a star of candidates (0 -> 1; 1..5 -> 0) plus the tour: every degree kept
    star = [(0, 1)] + [(a, 0) for a in range(1, 6)]
    order = [0, 1, 2, 3, 4, 5]
    offsets, neighbors = O.csr(star + [(order[i], order[(i + 1) % 6]) for i in range(6)], 6)
    degrees = np.diff(offsets)
    assert degrees.max() == 5 and offsets[-1] // 2 == 9
    rows = {a: set(neighbors[offsets[a]:offsets[a + 1]].tolist()) for a in range(6)}
    assert all((b in rows[a]) == (a in rows[b]) for a in range(6) for b in range(6))
    assert all(list(neighbors[offsets[a]:offsets[a + 1]]) == sorted(rows[a]) for a in range(6))
    # out-of-range cities are never members
    tour = np.array(order, dtype=np.uint32)
    assert not O.hard_two_opt_allowed(0, 1, 6, 2, (offsets, neighbors), 6)
    assert not O.hard_relocation_allowed(tour, 0, 0, (offsets, neighbors))


def test_select_next_exhausts_instead_of_leaving_the_graph():
    # the reference's regression: the graph is the 6-city tour itself; every
    # candidate's relocation after city 0 adds an edge outside it
    order = [0, 1, 2, 3, 4, 5]
    ring = O.csr(sorted(cycle_edges(order)), 6)
    xy = np.array([[0, 0], [1, 0], [2, 0], [3, 0], [4, 0], [5, 0]], dtype=np.int32)
    visited = np.array([1, 1, 0, 0, 0, 0], dtype=np.uint8)
    chosen, stage = O.select_next_hard(xy, [2, 3], [4, 5], [1.0, 1.0], visited, 0, 0.5, order, ring)
    assert stage == 4 and chosen == 0
    # with the chord (0, 2), (1, 3) the relocation of 2 after 0 stays inside
    chordal = O.csr(sorted(cycle_edges(order) | {(0, 2), (1, 3)}), 6)
    chosen, stage = O.select_next_hard(xy, [2, 3], [4, 5], [1.0, 1.0], visited, 0, 0.5, order, chordal)
    assert (chosen, stage) == (2, 0)


def _instances(n, count, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 10001, size=(count, n, 2)).astype(np.int32)


def test_hard_runs_stay_in_the_graph_and_equal_plain_faco_on_the_complete_graph():
    n, count, batches = 60, 2, 6
    xys = _instances(n, count, 17)
    s = O.default_settings(ants=16)
    keys, seeds = np.array([5, 6], dtype=np.uint64), np.array([1, 2, 3, 4], dtype=np.uint64)
    plain = O.run_faco(xys, keys, seeds, s, batches, 8)
    complete = [O.csr([(a, b) for a in range(n) for b in range(a + 1, n)], n) for _ in range(count)]
    same = O.run_faco(xys, keys, seeds, s, batches, 8, graphs=complete)
    assert same["tours"].tolist() == plain["tours"].tolist() and same["work"].tolist() == plain["work"].tolist()
    # the graph of the reference's from_candidates: primary candidates plus the initial tour
    graphs, members = [], []
    for g in range(count):
        prep = O.prepare(xys[g], s)
        tour = prep["initial_tour"].tolist()
        edges = {tuple(sorted((a, int(b)))) for a in range(n) for b in prep["primary"][a]}
        edges |= cycle_edges(tour)
        graphs.append(O.csr(sorted(edges), n))
        members.append(edges)
    hard = O.run_faco(xys, keys, seeds, s, batches, 8, graphs=graphs)
    for c in range(seeds.size):
        assert cycle_edges(hard["tours"][c].tolist()) <= members[c % count], c
    assert hard["work"].tolist() != plain["work"].tolist()     # the restriction changes the search
