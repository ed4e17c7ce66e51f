"""Escape primitives on the device against the oracle, bit for bit: the
escape random stream, slot views (construction and LS domains, with the LS
order by EUC_2D distance), and edge transactions (proposal, status, commit,
cache after). Plus the independent expectations of the reference's device
test (tests/cpp/escape_primitives.cu:59-84): capacity, undirected
de-duplication, all-or-nothing commits and expiry.
"""
import numpy as np
import pytest

from tests.conftest import announce

pytestmark = pytest.mark.gpu


def test_escape_streams_match_oracle(ext, oracle_mod):
    announce("escape-draws")
    for key, batch, ant, event, ls in [(991, 17, 3, 5, False), (2 ** 63 + 5, 0, 0, 0, True),
                                       (12345, 7, (1 << 30) - 1, 0xFFFFFFFF, False), (7, 1, 2, 3, True)]:
        got = ext.diagnostic_escape_draws(key, batch, ant, event, ls, 37)
        np.testing.assert_array_equal(got, oracle_mod.escape_draws(key, batch, ant, event, ls, 37))


def _rows(rng, rows, width, bw, n):
    original = np.full((rows, width), n, dtype=np.uint32)
    backup = np.full((rows, bw), n, dtype=np.uint32)
    from_cities = rng.integers(0, n, rows).astype(np.uint32)
    for r in range(rows):
        m = int(rng.integers(0, width + 1))
        pool = [v for v in range(n) if v != from_cities[r]]
        original[r, :m] = rng.choice(pool, m, replace=False)
        k = int(rng.integers(0, bw + 1))
        # backups: nearby ids with repeats, the row's own city and row members mixed in
        cand = rng.integers(0, n, k).astype(np.uint32)
        if k > 2:
            cand[rng.integers(0, k)] = from_cities[r]
            if m:
                cand[rng.integers(0, k)] = original[r, 0]
            cand[rng.integers(0, k)] = cand[0]
        backup[r, :k] = cand
        if k > 4 and rng.random() < 0.3:
            backup[r, k // 2] = n   # an early padding entry ends the list
    return original, backup, from_cities


@pytest.mark.parametrize("width,bw,ls", [(16, 64, False), (20, 64, True), (20, 10, True), (8, 3, False)])
def test_slot_views_match_oracle(ext, oracle_mod, width, bw, ls):
    rng = np.random.default_rng(width * 100 + bw)
    n, rows = 300, 200
    original, backup, from_cities = _rows(rng, rows, width, bw, n)
    xy = rng.integers(0, 10001, size=(n, 2)).astype(np.float64)
    events = rng.integers(0, 1 << 32, rows, dtype=np.uint64).astype(np.uint32)
    key, batch = 0x9E3779B97F4A7C15, 11
    announce("escape-rows", width, bw, ls)
    for sort in (False, True):
        got = ext.diagnostic_escape_rows(original, backup, n, from_cities, key, batch, events, ls=ls,
                                         coordinates=xy if sort else None)
        replaced_rows = 0
        for r in range(rows):
            dist = None
            if sort:
                d = np.rint(np.sqrt(((xy - xy[from_cities[r]]) ** 2).sum(axis=1)))   # EUC_2D of every city
                dist = np.append(d, np.inf)
            want = oracle_mod.escape_slot_view(original[r], backup[r], n, int(from_cities[r]), key=key, batch=batch,
                                               ant=r, event=int(events[r]), ls=ls, distance_of=dist)
            np.testing.assert_array_equal(got["members"][r], want["members"], err_msg=f"row {r} sort {sort}")
            np.testing.assert_array_equal(got["replaced"][r], want["replaced"], err_msg=f"row {r}")
            assert got["valid"][r] == want["valid"] and got["replacements"][r] == want["replacements"], r
            replaced_rows += want["replacements"] > 0
        assert replaced_rows > rows // 4


def _random_cases(rng, cases, n):
    graph = {tuple(sorted(rng.choice(n, 2, replace=False).tolist())) for _ in range(3 * n)}
    csr = _csr(graph, n)
    counts = rng.integers(1, 4, cases).astype(np.uint32)
    removed = rng.integers(0, n, size=(cases, 3, 2)).astype(np.uint32)
    added = rng.integers(0, n + 2, size=(cases, 3, 2)).astype(np.uint32)    # some invalid endpoints
    for c in range(cases):
        if rng.random() < 0.3:
            added[c, 0] = removed[c, 0][::-1]            # an edge put back unchanged
        if rng.random() < 0.2:
            added[c, 1] = added[c, 0][::-1]              # a duplicate within the group
    permits = rng.integers(0, 2, cases).astype(np.uint32)
    sizes = np.where(rng.random(cases) < 0.5, rng.integers(0, 65, cases), rng.integers(60, 65, cases)).astype(np.uint32)
    cache = np.zeros((cases, 64, 2), dtype=np.uint32)
    for c in range(cases):
        cache[c, : sizes[c]] = rng.integers(0, n, size=(sizes[c], 2))
        if sizes[c] and rng.random() < 0.4:
            cache[c, int(rng.integers(0, sizes[c]))] = added[c, 0][::-1]   # already in the cache
    epochs = rng.integers(1, 1 << 40, cases, dtype=np.uint64)
    return counts, removed, added, permits, cache, sizes, epochs, csr


def _csr(graph, n):
    rows = [[] for _ in range(n)]
    for a, b in graph:
        rows[a].append(b)
        rows[b].append(a)
    offsets = np.zeros(n + 1, dtype=np.uint32)
    neighbors = []
    for a in range(n):
        r = sorted(set(rows[a]))
        neighbors += r
        offsets[a + 1] = offsets[a] + len(r)
    return offsets, np.array(neighbors, dtype=np.uint32)


def test_transactions_match_oracle(ext, oracle_mod):
    rng = np.random.default_rng(3)
    n, cases = 30, 600
    counts, removed, added, permits, cache, sizes, epochs, csr = _random_cases(rng, cases, n)
    announce("escape-transactions", cases)
    got = ext.diagnostic_escape_transactions(counts, removed, added, permits, cache, sizes, epochs, *csr)
    statuses = set()
    for c in range(cases):
        k = int(counts[c])
        want = oracle_mod.escape_transaction(removed[c, :k], added[c, :k], csr, n, cache[c, : sizes[c]].tolist(),
                                             int(epochs[c]), bool(permits[c]))
        p = got["proposals"][c]
        assert oracle_mod.STATUS[int(p[0])] == want["status"], c
        assert int(p[1]) == want["count"] and int(p[2]) == want["expected_size"], c
        assert (int(p[3]) | (int(p[4]) << 32)) == want["expected_epoch"], c
        assert [(int(p[5 + 2 * i]), int(p[6 + 2 * i])) for i in range(want["count"])] == want["edges"], c
        assert bool(got["committed"][c]) == want["committed"], c
        assert got["cache_after"][c, : got["size_after"][c]].tolist() == [list(e) for e in want["cache_after"]], c
        statuses.add(want["status"])
    assert statuses == {"allowed", "invalid_edge", "outside_graph", "capacity"}


def test_reference_device_expectations(ext):
    """tests/cpp/escape_primitives.cu:17-84 restated: caches of 0, 61, 62, 63
    and 64 edges {0, j + 1}, a three-edge move with 3, 2, 1, 0 or 2 genuinely
    new edges, with and without permission, on an empty graph of 100 cities."""
    lengths, genuinely_new = [0, 61, 62, 63, 64], [3, 2, 1, 0, 2]
    removed = [[4, 5], [5, 6], [6, 7]]
    cases = 50
    counts = np.full(cases, 3, np.uint32)
    rem = np.zeros((cases, 3, 2), np.uint32)
    add = np.zeros((cases, 3, 2), np.uint32)
    permits = np.zeros(cases, np.uint32)
    cache = np.zeros((cases, 64, 2), np.uint32)
    sizes = np.zeros(cases, np.uint32)
    for i in range(cases):
        length, variant, permit = lengths[i // 10], (i // 2) % 5, i % 2
        a = [[70, 71], [72, 73], [74, 75]]
        if variant == 1:
            a[1] = [71, 70]
        if variant == 2:
            a[1] = a[2] = [71, 70]
        if variant == 3:
            a = [list(r) for r in removed]
        if variant == 4:
            a[0] = [5, 4]
        rem[i], add[i], permits[i], sizes[i] = removed, a, permit, length
        for j in range(length):
            cache[i, j] = [0, j + 1]
    offsets = np.zeros(101, np.uint32)
    announce("escape-reference-cases")
    got = ext.diagnostic_escape_transactions(counts, rem, add, permits, cache, sizes, np.ones(cases, np.uint64),
                                             offsets, np.zeros(0, np.uint32))
    for i in range(cases):
        count, length = genuinely_new[(i // 2) % 5], lengths[i // 10]
        expected = count == 0 or (i % 2 == 1 and length + count <= 64)
        status = 0 if expected else (3 if i % 2 else 2)
        p = got["proposals"][i]
        assert bool(got["committed"][i]) == expected and int(p[0]) == status, i
        assert int(got["size_after"][i]) == length + (count if expected else 0), i
        if expected or i % 2:
            assert int(p[1]) == count, i
        # the cache before the commit is read-only to the proposal
        assert got["cache_after"][i, :length].tolist() == [[0, j + 1] for j in range(length)], i
