"""The oracle's Or-opt extension of the checklist local search, pinned by
brute force (the reference has no Or-opt).

- one popped city: the oracle's chosen move equals the best move of an
  independent enumeration of the 2-opt + Or-opt neighbourhood in which every
  move is built as an edge set and its gain is the cost difference of the two
  edge sets (no gain formula), with the tie rule "first in scan order";
- whole runs: the cost the search maintains from move gains equals the cost
  recomputed from the final tour, the tour is a permutation, and Or-opt
  moves are actually taken;
- a hand case on eight cities that the search solves to the optimum.
"""
import numpy as np
import pytest

from oracle import oracle as O


def rows(M, width):
    n = M.shape[0]
    return np.array([sorted((b for b in range(n) if b != a), key=lambda b: (M[a, b], b))[:width]
                     for a in range(n)], dtype=np.uint32)


def edges_of(tour):
    n = len(tour)
    return {frozenset((int(tour[i]), int(tour[(i + 1) % n]))) for i in range(n)}


def cost_of(E, M):
    return sum(int(M[tuple(e)]) for e in E)


def neighbourhood(tour, a, ls_row, M):
    """Moves of city a in scan order: (gain, new edge set)."""
    n = len(tour)
    pos = {int(v): i for i, v in enumerate(tour)}
    succ = lambda v: int(tour[(pos[v] + 1) % n])  # noqa: E731
    pred = lambda v: int(tour[(pos[v] - 1) % n])  # noqa: E731
    E = edges_of(tour)
    base = cost_of(E, M)
    moves = []
    a_next, a_prev = succ(a), pred(a)
    passed = [0, 0]
    for kind in (0, 1):
        cur = M[a, a_next] if kind == 0 else M[a_prev, a]
        for b in ls_row:
            b = int(b)
            if not cur > M[a, b]:
                break
            passed[kind] += 1
            nb = succ(b) if kind == 0 else pred(b)
            closing = a_next if kind == 0 else a_prev
            if nb == a or b == closing:
                moves.append((0, E))
                continue
            E2 = (E - {frozenset((a, closing)), frozenset((b, nb))}) | {frozenset((a, b)), frozenset((closing, nb))}
            moves.append((base - cost_of(E2, M), E2))
    p = a_prev
    for j in range(passed[1]):
        b = int(ls_row[j])
        seg = [a]
        for L in (1, 2, 3):
            if L + 4 > n:
                break
            if L > 1:
                seg.append(succ(seg[-1]))
            last, nx = seg[-1], succ(seg[-1])
            for forward in (True, False):
                c, d = (b, succ(b)) if forward else (pred(b), b)
                if any(v in seg or v in (p, nx) for v in (c, d)):
                    continue
                new = {frozenset((p, nx))} | ({frozenset((b, a)), frozenset((last, d))} if forward
                                              else {frozenset((c, last)), frozenset((a, b))})
                E2 = (E - {frozenset((p, a)), frozenset((last, nx)), frozenset((c, d))}) | new
                moves.append((base - cost_of(E2, M), E2))
    return moves


@pytest.mark.parametrize("n", [5, 6, 8, 12, 25, 60])
def test_single_city_choice_matches_brute_force(n):
    rng = np.random.default_rng(n)
    taken = 0
    for rep in range(150):
        xy = rng.integers(0, [10001, 40][rep % 2], size=(n, 2))
        M = O.distance_matrix(xy)
        ls = rows(M, min(8, n - 1))
        tour = rng.permutation(n).astype(np.uint32)
        a = int(rng.integers(n))
        moves = neighbourhood(tour, a, ls[a], M)
        if not moves:
            continue
        best, chosen = -1, None
        for gain, E2 in moves:
            if gain > best:
                best, chosen = gain, E2
        # a budget of exactly this city's evaluations stops the search after it
        out, cost, evals = O.two_opt(xy, ls, tour, [a], len(moves), or_opt=True)
        assert evals == len(moves)
        expected = chosen if best > 0 else edges_of(tour)
        assert edges_of(out) == expected, (n, rep, a)
        assert cost == cost_of(expected, M)
        taken += best > 0 and len(expected - edges_of(tour)) == 3
    assert taken > 0, "no Or-opt move was the best move in any case"


@pytest.mark.parametrize("n", [6, 17, 80, 300])
def test_runs_keep_cost_and_take_or_moves(n):
    rng = np.random.default_rng(100 + n)
    differ = 0
    for rep in range(20):
        xy = rng.integers(0, 10001, size=(n, 2))
        M = O.distance_matrix(xy)
        ls = rows(M, min(20, n - 1))
        parent = rng.permutation(n)
        visit = rng.permutation(n)
        for limit in (7, 100000):
            plain = O.construct_ls(xy, ls, parent, visit, 8, limit)
            withor = O.construct_ls(xy, ls, parent, visit, 8, limit, or_opt=True)
            t = withor["tour"]
            assert sorted(t.tolist()) == list(range(n))
            assert withor["final_cost"] == int(M[t, np.roll(t, -1)].sum())
            assert withor["construction_tour"].tolist() == plain["construction_tour"].tolist()
            assert withor["counters"]["move_evaluations"] <= limit
            differ += withor["tour"].tolist() != plain["tour"].tolist()
    assert differ > 0


def test_hand_case_reaches_the_optimum():
    # Seven cities on a line and one off it; the start tour visits the
    # outlier between 4 and 5. With Or-opt the search ends at the optimum
    # (checked over all 5040 tours), which moves the outlier into the
    # closing edge.
    import itertools
    xy = np.array([[0, 0], [100, 0], [200, 0], [300, 0], [400, 0], [500, 0], [200, 100], [600, 0]])
    M = O.distance_matrix(xy)
    ls = rows(M, 7)
    tour = np.array([0, 1, 2, 3, 4, 6, 5, 7], dtype=np.uint32)
    out, cost, _ = O.two_opt(xy, ls, tour, list(range(8)), 100000, or_opt=True)
    best = min(int(M[0, t[0]]) + sum(int(M[t[i], t[i + 1]]) for i in range(6)) + int(M[t[-1], 0])
               for t in itertools.permutations(range(1, 8)))
    assert cost == cost_of(edges_of(out), M) == best
    _, plain_cost, _ = O.two_opt(xy, ls, tour, list(range(8)), 100000)
    assert plain_cost >= cost
