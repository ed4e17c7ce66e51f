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

Rejected moves (faco_cpu.cpp:180-183): the count of one city's scan equals
the evaluated moves whose tour after the move leaves the graph (edge sets
again). Where the reference checker oracle/_ref is built, the Hard 2-opt
(tour, cost, evaluations, rejections) and the Hard selection equal the
reference core's, given its own SparseUndirectedGraph as the predicate.
"""
import itertools

import numpy as np
import pytest

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
    # a rejection is an evaluated move, and the complete graph rejects nothing
    assert same["constraint_rejections"].tolist() == [0] * seeds.size
    assert all(0 < hard["constraint_rejections"][c] <= hard["work"][c][1] for c in range(seeds.size))


def _dist(xy, a, b):
    dx, dy = float(xy[a][0] - xy[b][0]), float(xy[a][1] - xy[b][1])
    return int((dx * dx + dy * dy) ** 0.5 + 0.5)


def test_hard_two_opt_rejections_are_the_moves_that_leave_the_graph():
    """One city's 2-opt scan (the budget ends the run after it): the moves it
    evaluates are the LS neighbours b nearer to a than a's successor (kind 0)
    or predecessor (kind 1), in row order; the oracle's rejection count is the
    number of them whose tour after the move (carried out on a list) has an
    edge outside the graph."""
    rng = np.random.default_rng(23)
    cases = rejected_total = 0
    for trial in range(60):
        n = int(rng.integers(8, 30))
        xy = rng.integers(0, 1000, size=(n, 2)).astype(np.int32)
        lw = min(8, n - 1)
        ls = np.array([[b for b in sorted(range(n), key=lambda b: (_dist(xy, a, b), b)) if b != a][:lw]
                       for a in range(n)], dtype=np.uint32)
        tour = [0, *rng.permutation(np.arange(1, n)).tolist()]
        a = int(rng.integers(0, n))
        i = tour.index(a)
        edges = cycle_edges(tour) | {tuple(sorted((u, int(v)))) for u in range(n) for v in ls[u]
                                     if rng.random() < 0.5}
        graph = O.csr(sorted(edges), n)
        evaluated = rejected = 0
        for kind in (0, 1):
            current = _dist(xy, a, tour[(i + 1) % n]) if kind == 0 else _dist(xy, tour[i - 1], a)
            for b in ls[a].tolist():
                if not current > _dist(xy, a, b):
                    break
                evaluated += 1
                rejected += not cycle_edges(two_opt_after(tour, a, b, kind)) <= edges
        if evaluated == 0:
            continue
        _, _, evals, got = O.two_opt_hard(xy, ls, tour, [a], evaluated, graph)
        assert (evals, got) == (evaluated, rejected), (trial, n, a)
        cases += 1
        rejected_total += rejected
    assert cases > 30 and rejected_total > 0


needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checker oracle/_ref not built here")


def _random_graph(rng, n, tour, ls, share):
    """The tour's edges, a share of the LS row edges and a few random chords."""
    edges = set(cycle_edges(tour))
    for a in range(n):
        for b in ls[a]:
            if rng.random() < share:
                edges.add(tuple(sorted((a, int(b)))))
    for _ in range(n // 4):
        a, b = (int(v) for v in rng.choice(n, 2, replace=False))
        edges.add(tuple(sorted((a, b))))
    return sorted(edges)


@needs_ref
def test_hard_two_opt_against_reference_core():
    """The oracle's Hard 2-opt equals the reference's checklist_two_opt given
    its own SparseUndirectedGraph as the edge predicate: tour, cost, move
    evaluations and constraint rejections, over random sparse graphs, tours,
    checklists and budgets (cut short and unbounded)."""
    rng = np.random.default_rng(31)
    rejected = limited = 0
    for trial in range(150):
        n = int(rng.integers(6, 60))
        xy = rng.integers(0, 2000 if trial % 3 else 12, size=(n, 2)).astype(np.int32)
        M = O.distance_matrix(xy)
        lw = min(int(rng.integers(3, 12)), n - 1)
        ls = np.array([sorted((b for b in range(n) if b != a), key=lambda b: (M[a, b], b))[:lw] for a in range(n)],
                      dtype=np.uint32)
        tour = rng.permutation(n).tolist()
        edges = _random_graph(rng, n, tour, ls, float(rng.choice([0.0, 0.3, 0.7])))
        checklist = rng.permutation(n)[:int(rng.integers(1, n + 1))]
        limit = int(rng.choice([3, 17, 2**64 - 1]))
        got_tour, got_cost, evals, rejections = O.two_opt_hard(xy, ls, tour, checklist, limit, O.csr(edges, n))
        ref_tour, ref_cost, counters = O.ref_two_opt_hard(M.astype(np.float64), ls, tour, checklist, limit, edges)
        assert got_tour.tolist() == ref_tour.tolist(), trial
        assert got_cost == ref_cost and evals == counters[2] and rejections == counters[5], trial
        assert cycle_edges(got_tour.tolist()) <= set(edges) | cycle_edges(tour)
        rejected += rejections
        limited += counters[6]
    assert rejected > 100 and limited > 10


@needs_ref
def test_hard_select_next_against_reference_core():
    """The oracle's Hard select_next equals the reference's select_next with
    the relocation predicate as its node filter (every stage, exhaustion
    included)."""
    rng = np.random.default_rng(37)
    stages = set()
    for k in range(400):
        n = int(rng.integers(5, 24))
        xy = rng.integers(0, 60, size=(n, 2)).astype(np.int32)
        M = O.distance_matrix(xy)
        tour = rng.permutation(n).tolist()
        cur = int(rng.integers(0, n))
        order = sorted((b for b in range(n) if b != cur), key=lambda b: (M[cur, b], b))
        pw = min(int(rng.integers(2, 8)), n - 1)
        bw = min(int(rng.integers(0, 5)), n - 1 - pw)
        ls = np.array([sorted((b for b in range(n) if b != a), key=lambda b: (M[a, b], b))[:pw] for a in range(n)])
        edges = _random_graph(rng, n, tour, ls, float(rng.choice([0.0, 0.2, 0.6, 1.0])))
        visited = (rng.random(n) < rng.random()).astype(np.uint8)
        visited[cur] = 1
        products = rng.random(pw) * (k % 5 != 0)
        u = float(rng.random())
        got = O.select_next_hard(xy, order[:pw], order[pw:pw + bw], products, visited, cur, u, tour, O.csr(edges, n))
        want = O.ref_select_next_hard(M.astype(np.float64), order[:pw], order[pw:pw + bw], products, visited, cur, u,
                                      tour, edges)
        assert got == want, k
        stages.add(got[1])
    assert stages == {0, 1, 2, 3, 4}


@needs_ref
def test_padded_ls_rows_against_reference_core():
    """A graph spec's LS rows are right-padded with n: the oracle's Hard 2-opt
    over padded rows equals the reference core over the same rows without
    the padding (a padded slot ends the row, faco_device.cuh:320)."""
    rng = np.random.default_rng(41)
    shorter = 0
    for trial in range(80):
        n = int(rng.integers(8, 50))
        xy = rng.integers(0, 2000, size=(n, 2)).astype(np.int32)
        M = O.distance_matrix(xy)
        tour = rng.permutation(n).tolist()
        knn = np.array([sorted((b for b in range(n) if b != a), key=lambda b: (M[a, b], b))[:6] for a in range(n)])
        edges = _random_graph(rng, n, tour, knn, 0.5)
        nbrs = [[] for _ in range(n)]
        for a, b in edges:
            nbrs[a].append(b)
            nbrs[b].append(a)
        lw = 8
        ls = np.full((n, lw), n, dtype=np.uint32)
        for a in range(n):
            row = sorted(nbrs[a], key=lambda b: (M[a, b], b))[:lw]
            ls[a, :len(row)] = row
            shorter += len(row) < lw
        checklist = rng.permutation(n)[:int(rng.integers(1, n + 1))]
        limit = int(rng.choice([5, 2**64 - 1]))
        got_tour, got_cost, evals, rejections = O.two_opt_hard(xy, ls, tour, checklist, limit, O.csr(edges, n))
        ref_tour, ref_cost, counters = O.ref_two_opt_hard(M.astype(np.float64), ls, tour, checklist, limit, edges)
        assert got_tour.tolist() == ref_tour.tolist(), trial
        assert (got_cost, evals, rejections) == (ref_cost, counters[2], counters[5]), trial
    assert shorter > 100


def test_spec_rows_equal_to_the_prepared_ones_change_nothing():
    n, count, batches = 50, 2, 4
    xys = _instances(n, count, 29)
    s = O.default_settings(ants=16, backup_width=6)
    keys, seeds = np.array([8, 9], dtype=np.uint64), np.array([1, 2, 3, 4], dtype=np.uint64)
    complete = [O.csr([(a, b) for a in range(n) for b in range(a + 1, n)], n) for _ in range(count)]
    plain = O.run_faco(xys, keys, seeds, s, batches, 8, graphs=complete)
    rows = []
    for g in range(count):
        prep = O.prepare(xys[g], s)
        rows.append((prep["primary"], prep["backup"], prep["ls"]))
    same = O.run_faco(xys, keys, seeds, s, batches, 8, graphs=complete, rows=rows)
    assert same["tours"].tolist() == plain["tours"].tolist() and same["work"].tolist() == plain["work"].tolist()
