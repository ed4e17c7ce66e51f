"""The oracle's escape primitives, pinned by the reference's own tests:

- slot views: the 84 scripted slot-rounding cases of
  tests/cpp/test_escape_slots.cpp:19-54 (valid members, per-slot origin
  flags, the LS order by (distance, id), no invented candidates when backups
  run out, and the expected quarter of replaced slots over the four
  roundings);
- edge transactions: the exhaustive move check of
  tests/cpp/test_escape_edges.cpp:29-88 (every relocation and 2-opt move of
  every tour with n = 3..6, sampled at n = 7, under empty and seeded caches,
  with and without permission) against the complete before/after tour edge
  sets, and its capacity / expiry cases (:90-121);
- random streams: domain separation and replay as in
  tests/cpp/escape_primitives.cu:42-56,77-81.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def test_slot_rounding_cases():
    backup = [0, 1, 51, 51, 52, 53, 54, 55, 56, 100]
    distance = np.array([node % 7 for node in range(101)], dtype=np.float64)
    cases = 0
    for m in range(21):
        original = [100] * 20
        for j in range(m):
            original[j] = j + 1
        total = 0
        for rounding in range(4):
            row = O.escape_slot_view(original, backup, 100, 0, script=rounding)
            seen, count = set(), 0
            for j in range(20):
                member, flag = int(row["members"][j]), bool(row["replaced"][j])
                if j >= m:
                    assert member == 100 and not flag, "padding slot became a member"
                else:
                    assert 0 < member < 100 and member not in seen, "self, duplicate or invalid member"
                    seen.add(member)
                    replacement = member not in original
                    assert flag == replacement, "per-slot origin flag"
                    count += replacement
            total += count
            ls = O.escape_slot_view(original, backup, 100, 0, script=rounding, distance_of=distance)
            for j in range(1, m):
                prev, cur = int(ls["members"][j - 1]), int(ls["members"][j])
                assert (distance[prev], prev) < (distance[cur], cur), "LS order by (distance, id)"
            for j in range(m):
                assert bool(ls["replaced"][j]) == (int(ls["members"][j]) not in original), "flag lost in the sort"
            short = O.escape_slot_view(original, [0, 1, 1, 100], 100, 0, script=rounding)
            assert short["replacements"] <= 1 and (m == 0 or short["replacements"] == 0), "invented candidate"
            cases += 1
        assert total == m, "the four equally likely roundings must replace m/4 slots on average"
    assert cases == 84


def _edges(tour):
    n = len(tour)
    return {frozenset((int(tour[i]), int(tour[(i + 1) % n]))) for i in range(n)}


def _relocate(tour, target, node):
    t = [v for v in tour if v != node]
    t.insert(t.index(target) + 1, node)
    return t


def _reverse_section(tour, start, end):
    # CpuTour::reverse_section (cpp/src/faco_cpu.cpp:137-146): reverse the
    # positions [pos(start), pos(end)) after ordering the two endpoints
    t = list(tour)
    i, j = t.index(start), t.index(end)
    if i > j:
        i, j = j, i
    t[i:j] = reversed(t[i:j])
    return t


def _verify(before, after, graph_set, csr, n, prepare, counters):
    old, now = _edges(before), _edges(after)
    for seeded in (False, True):
        for permit in (False, True):
            cache = []
            if seeded:
                cache = [(a, b) for a in range(n) for b in range(a + 1, n)
                         if frozenset((a, b)) not in graph_set and (a + b) % 2 == 0]
            cache_set = {frozenset(e) for e in cache}
            novel = {e for e in now if e not in old and e not in graph_set and e not in cache_set}
            expected = not novel or (permit and len(cache) + len(novel) <= 64)
            prop = prepare(cache, permit)
            assert (prop["status"] == "allowed") == expected, (before, after, seeded, permit, prop)
            if "committed" in prop:
                assert prop["committed"] == expected
                if expected:
                    assert prop["count"] == len(novel)
                    assert {frozenset(e) for e in prop["cache_after"]} == cache_set | novel
                    assert len(prop["cache_after"]) == len(cache) + len(novel)
                    counters["permitted"] += 1
                else:
                    assert prop["cache_after"] == [tuple(e) for e in cache]
                    counters["rejected"] += 1
            elif expected:
                assert prop["count"] == len(novel)
            inherited = [e for e in now if e in old and e not in graph_set and e not in cache_set]
            counters["inherited"] += bool(inherited) and expected


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7])
def test_transactions_against_tour_edge_sets(n):
    rng = np.random.default_rng(n)
    graph = [(i, (i + 1) % n) for i in range(n)]
    graph_set = {frozenset(e) for e in graph}
    csr = O.csr(graph, n)
    counters = {"permitted": 0, "rejected": 0, "inherited": 0}
    perms = [list(p) for p in itertools.permutations(range(1, n))]
    if n == 7:
        perms = [perms[i] for i in rng.choice(len(perms), 40, replace=False)]
    for rest in perms:
        order = [0] + rest
        succ = {order[i]: order[(i + 1) % n] for i in range(n)}
        pred = {order[i]: order[i - 1] for i in range(n)}
        for a in range(n):
            for b in range(n):
                if a == b:
                    continue
                after = _relocate(order, a, b)
                _verify(order, after, graph_set, csr, n,
                        lambda cache, permit: O.escape_relocation(order, a, b, csr, cache, 1, permit), counters)
                for kind in (0, 1):
                    an = succ[a] if kind == 0 else pred[a]
                    bn = succ[b] if kind == 0 else pred[b]
                    after = _reverse_section(order, an if kind == 0 else a, bn if kind == 0 else b)
                    _verify(order, after, graph_set, csr, n,
                            lambda cache, permit: O.escape_transaction([(a, an), (b, bn)], [(a, b), (an, bn)], csr, n,
                                                                       cache, 1, permit), counters)
    assert counters["permitted"]
    if n >= 4:   # a triangle's graph holds every edge: nothing to reject
        assert counters["rejected"]
    if n >= 5:
        assert counters["inherited"], "no move kept an out-of-graph parent edge"


def test_capacity_and_expiry():
    csr = O.csr([], 20)
    cache = [(a, b) for a in range(20) for b in range(a + 1, 20)][:62]
    removed = [(8, 9), (9, 10), (10, 11)]
    prop = O.escape_transaction(removed, [(12, 13), (14, 15), (16, 17)], csr, 20, cache, 1, True)
    assert prop["status"] == "capacity" and not prop["committed"] and prop["cache_after"] == cache
    prop = O.escape_transaction(removed, [(12, 13), (13, 12), (14, 15)], csr, 20, cache, 1, True)
    assert prop["count"] == 2 and prop["committed"] and len(prop["cache_after"]) == 64
    full = prop["cache_after"]
    again, after = O.escape_commit(prop, full, 1)       # length expired: no double commit
    assert not again and after == full
    keep = O.escape_transaction(removed, [(9, 8), (9, 10), (10, 11)], csr, 20, full, 1, False)
    assert keep["status"] == "allowed" and keep["count"] == 0
    reset, after = O.escape_commit(keep, [], 2)          # the batch reset the cache: epoch expired
    assert not reset and after == []
    outside = O.escape_transaction([(1, 2)], [(12, 13)], csr, 20, [], 2, False)
    assert outside["status"] == "outside_graph", "last batch's exception became permanent"
    invalid = O.escape_transaction([(1, 2)], [(20, 3)], csr, 20, [], 2, True)
    assert invalid["status"] == "invalid_edge"


def test_escape_streams_are_separated_and_replayable():
    key = 991
    prefixes = set()
    for ant in range(25):
        for domain in range(6):
            if domain == 0:   # the ordinary ant stream of step 5
                words = O.philox([5, 0, ant, 17], [key & 0xFFFFFFFF, key >> 32])
            else:
                words = O.escape_draws(key, 18 if domain == 5 else 17, ant + (100 if domain == 4 else 0),
                                       6 if domain == 3 else 5, domain == 2, 4)
                again = O.escape_draws(key, 18 if domain == 5 else 17, ant + (100 if domain == 4 else 0),
                                       6 if domain == 3 else 5, domain == 2, 4)
                assert np.array_equal(words, again), "escape stream not replayable"
            prefix = tuple(int(w) for w in words)
            assert prefix not in prefixes, "streams declared separate collided"
            prefixes.add(prefix)
    # a stream is one word sequence: longer reads extend shorter ones
    long = O.escape_draws(key, 3, 4, 5, True, 37)
    assert np.array_equal(long[:9], O.escape_draws(key, 3, 4, 5, True, 9))
