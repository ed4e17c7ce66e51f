"""The oracle's Escape mode, pinned the way the reference pins its own
(tests/cpp/test_escape_engine.cpp:1-273):

- with the escape disabled an Escape run is the Hard run, tour for tour and
  counter for counter (the reference's `same(hard, disabled)`), plain and
  controlled;
- every ant's accepted moves, replayed from the parent tour on the
  reference's array tour (CpuTour::relocate / reverse_section), rebuild its
  tour; a move that adds an edge outside the graph and outside the ant's
  cache is a permitted one, the cache grows by exactly those edges (at most
  64, the sizes logged per move), relocations that open new edges come from
  a backup row city in a primary slot, 2-opt moves through replaced slots
  only at anchors; the anchors are the five endpoints of each permitted
  relocation; every LS row keeps its valid count, replaced members come from
  the backup row, rows are ordered by (distance, city), at most a quarter
  (rounded up) of a row is replaced, the footprint marks the rows with
  replacements, and a city that is not an anchor keeps its LS row;
- coverage as the reference demands: new edges in construction and in the
  2-opt, reactivations by the parent's footprint, footprints on the complete
  graph.
The graphs are the reference test's: the complete graph on a degenerate
instance (all cities at one point but city 0) and the initial tour plus
each city's four nearest, rows sorted by (distance, city), backup rows the
remaining cities in id order.
"""
import numpy as np
import pytest

from oracle import oracle as O

N_COUNTERS = len(O.ESCAPE_COUNTERS)
IDX = {name: i for i, name in enumerate(O.ESCAPE_COUNTERS)}


def euc(xy, a, b):
    return int(np.floor(np.hypot(*(xy[a] - xy[b]).astype(np.float64)) + 0.5))


def reference_problem(n, full, variant=0, ants=4):
    """The reference test's instances and graph specs (test_escape_engine.cpp:25-54)
    on integer coordinates: the fractional coordinates scaled by 30000."""
    if full:
        xy = np.zeros((n, 2), dtype=np.int64)
        xy[0, 0] = 1
    else:
        i = np.arange(n, dtype=np.int64)
        xy = np.stack([((i * 7919 + variant * 313) % 104729) * 30000 // 104729,
                       ((i * i * 13 + variant * 101 + 17) % 130363) * 30000 // 130363], axis=1)
    s = O.default_settings(ants=ants)
    prep = O.prepare(xy.astype(np.int32), s)
    norm = prep["settings"]
    pw, bw, lw = norm.primary_width, norm.backup_width, norm.ls_width
    tour = [int(v) for v in prep["initial_tour"]]
    members = {tuple(sorted((tour[k - 1], tour[k]))) for k in range(n)}
    if full:
        members |= {(a, b) for a in range(n) for b in range(a + 1, n)}
    else:
        for a in range(n):
            for b in prep["primary"][a][:4]:
                members.add(tuple(sorted((a, int(b)))))
    primary = np.full((n, pw), n, dtype=np.int64)
    ls = np.full((n, lw), n, dtype=np.int64)
    backup = np.full((n, bw), n, dtype=np.int64)
    for a in range(n):
        row = sorted((b for b in range(n) if b != a and tuple(sorted((a, b))) in members),
                     key=lambda b: (euc(xy, a, b), b))
        primary[a, :min(pw, len(row))] = row[:pw]
        ls[a, :min(lw, len(row))] = row[:lw]
        rest = [b for b in range(n) if b != a and b not in set(primary[a])][:bw]
        backup[a, :len(rest)] = rest
    return xy, s, norm, tour, members, (primary, backup, ls)


class ArrayTour:
    """The reference's CpuTour array moves (cpp/src/faco_cpu.cpp:90-146)."""

    def __init__(self, order):
        self.order = list(order)
        self.n = len(order)
        self.pos = [0] * self.n
        for i, v in enumerate(self.order):
            self.pos[v] = i

    def succ(self, v):
        return self.order[(self.pos[v] + 1) % self.n]

    def pred(self, v):
        return self.order[(self.pos[v] - 1) % self.n]

    def edges(self):
        return {tuple(sorted((self.order[k - 1], self.order[k]))) for k in range(self.n)}

    def relocate(self, target, node):
        if self.succ(target) == node:
            return
        pn, pt = self.pos[node], self.pos[target]
        if pt < pn:
            self.order[pt + 2:pn + 1] = self.order[pt + 1:pn]
            self.order[pt + 1] = node
        else:
            self.order[pn:pt] = self.order[pn + 1:pt + 1]
            self.order[pt] = node
        for k in range(min(pn, pt), max(pn, pt) + 1):
            self.pos[self.order[k]] = k

    def reverse_section(self, start, end):
        if start == end:
            return
        first, last = self.pos[start], self.pos[end]
        if first > last:
            first, last = last, first
        length = last - first
        if length <= self.n - length or first == 0:
            self.order[first:last] = self.order[first:last][::-1]
            for k in range(first, last):
                self.pos[self.order[k]] = k
        else:
            left, right = last, first - 1
            for _ in range((self.n - length) // 2):
                self.order[left], self.order[right] = self.order[right], self.order[left]
                self.pos[self.order[left]] = left
                self.pos[self.order[right]] = right
                left = (left + 1) % self.n
                right = (right - 1) % self.n


def verify_ant(xy, out, parent, rows, members, parent_fp, n, pw, lw, tally):
    """The reference's per-ant replay (test_escape_engine.cpp:147-219)."""
    primary, backup, ls = rows
    tour = ArrayTour(parent)
    cache = set()
    anchors = np.zeros(n, dtype=np.uint8)
    seen_ls = False
    for kind, a, b, slot, size, permit in out["events"].tolist():
        before = tour.edges()
        if kind == 0:
            assert not seen_ls and tour.succ(a) != b
            if permit:
                assert slot < pw and b in backup[a] and b not in primary[a]
                for v in (a, tour.succ(a), b, tour.pred(b), tour.succ(b)):
                    anchors[v] = 1
            tour.relocate(a, b)
        else:
            seen_ls = True
            assert slot < lw and out["ls_rows"][a, slot] == b and out["replaced"][a, slot] == permit
            assert not permit or anchors[a]
            tour.reverse_section(tour.succ(a) if kind == 1 else a, tour.succ(b) if kind == 1 else b)
        for e in tour.edges() - before:
            if e not in members and e not in cache:
                assert permit, "an unpermitted move added an edge outside the graph"
                cache.add(e)
                tally["new_edges"] += 1
                tally["ls_new_edges"] += kind != 0
        assert len(cache) <= 64 and size == len(cache)
    assert tour.order == out["tour"].tolist()
    cost = sum(euc(xy, tour.order[k - 1], tour.order[k]) for k in range(n))
    assert cost == out["costs"][1]
    assert out["anchors"].tolist() == anchors.tolist()
    recorded = [tuple(e) for e in out["cache"].tolist()]
    assert all(a < b for a, b in recorded) and len(set(recorded)) == len(recorded) == len(cache)
    assert set(recorded) == cache and out["counters"][IDX["new_edges"]] == len(cache)
    assert out["counters"][IDX["ls_anchor_nodes"]] == int(anchors.sum())
    for node in range(n):
        row, flags = out["ls_rows"][node], out["replaced"][node]
        valid = int((ls[node] < n).sum())
        replacements, seen = 0, set()
        for j in range(lw):
            if j >= valid:
                assert row[j] == n and not flags[j]
                continue
            assert row[j] < n and row[j] != node and row[j] not in seen
            seen.add(int(row[j]))
            replaced = row[j] not in ls[node]
            assert bool(flags[j]) == replaced and (not replaced or (anchors[node] and row[j] in backup[node]))
            replacements += replaced
            if j:
                assert (euc(xy, node, row[j - 1]), row[j - 1]) < (euc(xy, node, row[j]), row[j])
        assert replacements <= (valid + 3) // 4 and out["footprint"][node] == (replacements != 0)
        if not anchors[node]:
            assert row.tolist() == ls[node].tolist()
        tally["footprints"] += replacements != 0
    # the reactivations: every anchor and footprint city entered the checklist
    fresh = out["counters"][IDX["old_view_reactivations"]] + out["counters"][IDX["anchor_reactivations"]]
    assert fresh <= int(((anchors != 0) | (parent_fp != 0)).sum())
    return tally


@pytest.mark.parametrize("n,full", [(31, True), (31, False), (200, False)])
def test_escape_ants_replay_like_the_reference(n, full):
    xy, s, norm, tour, members, rows = reference_problem(n, full)
    pw, lw = norm.primary_width, norm.ls_width
    graph = O.csr(sorted(members), n)
    rng = np.random.default_rng(n + full)
    tally = {"new_edges": 0, "ls_new_edges": 0, "footprints": 0, "events": 0}
    for trial in range(24 if n <= 31 else 12):
        parent = list(rng.permutation(n)) if trial % 3 == 2 else tour
        parent_fp = (rng.random(n) < 0.1).astype(np.uint8)
        products = rng.random((n, pw)) * (rows[0] < n)
        start = int(rng.integers(n))
        out = O.escape_ant(xy, s, graph, parent, products, float(rng.random()), parent_fp, key=1000 + trial,
                           batch=trial % 5, ant=trial, start=start, target=64, rows=rows)
        tally["events"] += len(out["events"])
        verify_ant(xy, out, parent, rows, members, parent_fp, n, pw, lw, tally)
    assert tally["events"] > 0 and tally["footprints"] > 0
    if not full:
        assert tally["new_edges"] > 0 and tally["ls_new_edges"] > 0


def test_disabled_escape_ant_is_the_hard_ant():
    """With the gates disabled no slot is replaced, no edge leaves the graph
    and the ant's moves are the Hard ant's (the same products, stream and
    graph), whatever the parent footprint says besides its reactivations."""
    n = 200
    xy, s, norm, tour, members, rows = reference_problem(n, False)
    graph = O.csr(sorted(members), n)
    rng = np.random.default_rng(3)
    products = rng.random((n, norm.primary_width)) * (rows[0] < n)
    for ant in range(6):
        out = O.escape_ant(xy, s, graph, tour, products, 0.5, np.zeros(n, np.uint8), key=77, batch=1, ant=ant,
                           start=ant * 7, target=32, rows=rows, disabled=True)
        counters = dict(zip(O.ESCAPE_COUNTERS, out["counters"].tolist()))
        assert counters["construction_opportunities"] > 0
        assert all(v == 0 for k, v in counters.items() if k != "construction_opportunities")
        assert len(out["cache"]) == 0 and not out["anchors"].any() and not out["replaced"].any()
        assert {tuple(sorted((out["tour"][k - 1], out["tour"][k]))) for k in range(n)} <= members


@pytest.mark.parametrize("n,full", [(31, True), (31, False), (300, False)])
def test_disabled_escape_runs_are_hard_runs(n, full):
    """The reference's same(hard, disabled) check: plain and controlled
    runs (a restarting rule baseline) give the same tours, costs, work and
    rejections; only the construction opportunities are counted."""
    xy, s, norm, tour, members, rows = reference_problem(n, full)
    graph = O.csr(sorted(members), n)
    seeds = np.array([17, 29], dtype=np.uint64)
    xys = xy[None].astype(np.int32)
    hard = O.run_faco(xys, [1], seeds, s, 6, 8, graphs=[graph], rows=[rows])
    off = O.run_faco(xys, [1], seeds, s, 6, 8, graphs=[graph], rows=[rows], escape="disabled")
    for key in ("tours", "costs", "work", "constraint_rejections"):
        np.testing.assert_array_equal(hard[key], off[key])
    assert not off["escape_counters"][:, 1:].any() and (off["escape_counters"][:, 0] >= off["work"][:, 0]).all()
    policy = O.Baseline()
    policy.mne_level = policy.max_mne_level = 3
    policy.restart_mode, policy.restart_probability = 2, 1.0
    keys = np.array([1], dtype=np.uint64)
    hard = O.run_controlled(xys, keys, seeds, s, 4 * 8, "baseline", baseline=policy, graphs=[graph], rows=[rows])
    off = O.run_controlled(xys, keys, seeds, s, 4 * 8, "baseline", baseline=policy, graphs=[graph], rows=[rows],
                           escape="disabled")
    for key in ("tours", "costs", "work", "constraint_rejections"):
        np.testing.assert_array_equal(hard[key], off[key])
    assert [st["restarts"] for st in hard["states"]] == [st["restarts"] for st in off["states"]]


def test_escape_runs_reach_every_path():
    """The reference test's coverage demands: new edges, reactivations by
    the parent's footprint after restarts and selections, anchors, and the
    tours stay Hamiltonian; the escape stream leaves the ordinary one alone
    (a restart every batch keeps the restart count of the Hard run)."""
    n = 200
    xy, s, norm, tour, members, rows = reference_problem(n, False, ants=8)
    graph = O.csr(sorted(members), n)
    seeds = np.array([3, 5, 7, 11], dtype=np.uint64)
    xys = xy[None].astype(np.int32)
    plain = O.run_faco(xys, [1], seeds, s, 12, 8, graphs=[graph], rows=[rows], escape="on")
    policy = O.Baseline()
    policy.mne_level = policy.max_mne_level = 3
    policy.restart_mode, policy.restart_probability = 2, 1.0
    ctl = O.run_controlled(xys, np.array([1], dtype=np.uint64), seeds, s, 8 * 12, "baseline", baseline=policy,
                           graphs=[graph], rows=[rows], escape="on")
    hard = O.run_controlled(xys, np.array([1], dtype=np.uint64), seeds, s, 8 * 12, "baseline", baseline=policy,
                            graphs=[graph], rows=[rows])
    assert [st["restarts"] for st in ctl["states"]] == [st["restarts"] for st in hard["states"]]
    for run in (plain, ctl):
        total = dict(zip(O.ESCAPE_COUNTERS, run["escape_counters"].sum(axis=0).tolist()))
        for name in ("construction_gates", "construction_replaced_slots", "escape_relocations", "ls_anchor_nodes",
                     "ls_replaced_slots", "anchor_reactivations", "new_edges", "old_view_reactivations"):
            assert total[name] > 0, name
        # one gate per 16 opportunities on average
        assert 0.03 < total["construction_gates"] / total["construction_opportunities"] < 0.1
        for t in run["tours"]:
            assert sorted(t.tolist()) == list(range(n))
