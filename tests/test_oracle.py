"""The CPU oracle pinned before it is trusted (no GPU needed).

- golden vectors written from the reference's own CPU core
  (tests/golden/make_golden.py; reference cpp/src/faco_cpu.cpp,
  cpp/src/program_cpu.cpp);
- the live reference checker oracle/_ref where it is built;
- closed forms and published vectors: EUC_2D against the FP64 TSPLIB rule,
  Philox4x32-10 against Random123 known answers, the uniform53 word layout,
  unbiased rejection for start cities, SplitMix64 keys.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    return json.loads((GOLDEN / name).read_text())


# ---------------------------------------------------------------------------
# closed forms

def test_euc2d_matches_tsplib_fp64_rule():
    rng = np.random.default_rng(3)
    pairs = rng.integers(0, 32768, size=(20000, 4))
    for x1, y1, x2, y2 in pairs[:4000]:
        want = math.floor(math.sqrt(float((x1 - x2) ** 2 + (y1 - y2) ** 2)) + 0.5)
        assert O.euc2d(int(x1), int(y1), int(x2), int(y2)) == want
    # exhaustive small squares and every rounding boundary s = k^2 + k, k^2 + k + 1
    for k in range(0, 46341, 7):
        for s in (k * k, k * k + k, k * k + k + 1):
            dx = int(math.isqrt(s))
            dy2 = s - dx * dx
            dy = math.isqrt(dy2)
            if dy * dy != dy2 or dx > 32767 or dy > 32767:
                continue
            assert O.euc2d(0, 0, dx, dy) == math.floor(math.sqrt(float(s)) + 0.5)
    assert O.euc2d(0, 0, 3, 4) == 5
    assert O.euc2d(0, 0, 1, 1) == 1          # sqrt(2) = 1.414
    assert O.euc2d(0, 0, 1, 2) == 2          # sqrt(5) = 2.236
    assert O.euc2d(0, 0, 32767, 32767) == 46340     # sqrt(2) * 32767 = 46339.54
    M = O.distance_matrix(pairs[:300, :2])
    for a in range(0, 300, 17):
        for b in range(0, 300, 13):
            assert M[a, b] == O.euc2d(*map(int, pairs[a, :2]), *map(int, pairs[b, :2]))


def test_philox_known_answers():
    # Random123 philox4x32_10 known-answer vectors
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for counter, key, want in kat:
        assert tuple(int(w) for w in O.philox(counter, key)) == want


def test_uniform53_word_layout_and_domains():
    for key, batch, ant, step in [(1, 2, 3, 4), (2**64 - 1, 0, 0xFFFFFFFF, 0), (12345, 7, 0, 0xFFFFFFFE)]:
        w = O.philox([step, 0, ant, batch], [key & 0xFFFFFFFF, key >> 32])
        want = (((int(w[0]) << 32) | int(w[1])) >> 11) * 2.0 ** -53
        assert O.uniform53(key, batch, ant, step) == want
        assert 0.0 <= want < 1.0
    # the colony coin (ant = UINT32_MAX, step 0) and ant streams are distinct domains
    values = {O.uniform53(99, b, a, s) for b in range(3) for a in (0, 1, 0xFFFFFFFF) for s in (0, 1, 2)}
    assert len(values) == 27


def test_start_node_rejection_sampling():
    # first word accepted unless below 2^32 mod count; words continue in stream order
    for key in range(50):
        for count in (3, 1000, 10000):
            threshold = (2**32 - count) % count
            words = []
            ctr = [0xFFFFFFFF, 0, 5, 9]
            for _ in range(3):
                words += [int(x) for x in O.philox(ctr, [key, 0])]
                v = (ctr[0] | ctr[1] << 32 | ctr[2] << 64 | ctr[3] << 96) + 1
                ctr = [(v >> (32 * i)) & 0xFFFFFFFF for i in range(4)]
            want = next(x % count for x in words if x >= threshold)
            assert O.start_node(key, 9, 5, count) == want
    # heavy rejection: count just above 2^31 rejects about half of the words
    hits = [O.start_node(k, 0, 0, 2**31 + 1) for k in range(200)]
    assert all(0 <= h <= 2**31 for h in hits)


def test_colony_key_is_splitmix_of_seed_and_instance():
    mask = (1 << 64) - 1

    def mix(z):
        z = (z + 0x9E3779B97F4A7C15) & mask
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        return z ^ (z >> 31)
    # SplitMix64 of 0 from the finaliser's published first output
    assert mix(0) == 0xE220A8397B1DCDAF
    for key, seed in [(0, 0), (1000, 7), (2**63, 2**64 - 1)]:
        assert O.colony_key(key, seed) == mix(seed ^ mix((key + 0xD1B54A32D192ED03) & mask))


# ---------------------------------------------------------------------------
# golden vectors from the reference core

def test_construct_ls_golden():
    cases = load("construct_ls.json")
    assert len(cases) >= 100
    for case in cases:
        got = O.construct_ls(np.array(case["xy"]), np.array(case["ls"]), case["parent"], case["visit"],
                             case["mne_target"], case["limit"])
        assert got["construction_tour"].tolist() == case["construction_tour"]
        assert got["tour"].tolist() == case["tour"]
        assert got["positions"].tolist() == case["positions"]
        assert got["checklist"].tolist() == case["checklist"]
        assert got["counters"] == case["counters"]
        assert got["construction_cost"] == case["construction_cost"]
        assert got["final_cost"] == case["final_cost"]


def test_identity_move_three_city_regression():
    # the only 2-opt moves on a 3-city tour re-add the removed edges: no gain
    xy = np.array([[0, 0], [2500, 0], [2500, 1000]])
    ls = np.array([[1, 2], [2, 0], [1, 0]])
    got = O.construct_ls(xy, ls, [0, 1, 2], [0, 1, 2], 2, 1000)
    assert got["counters"]["accepted_moves"] == 0 and got["tour"].tolist() == [0, 1, 2]


def test_select_next_golden():
    cases = load("select_next.json")
    stages = set()
    for case in cases:
        chosen, stage = O.select_next(np.array(case["xy"]), case["primary"], case["backup"], case["products"],
                                      case["visited"], case["current"], case["uniform"])
        assert (chosen, stage) == (case["chosen"], case["stage"])
        stages.add(stage)
    assert {0, 1, 2, 3} <= stages  # primary, zero-weight, backup and global fallback all covered


def test_trail_limits_golden():
    for case in load("trail_limits.json"):
        lo, hi = O.trail_limits(case["width"], case["p_best"], case["retention"], case["cost"])
        assert lo == float.fromhex(case["minimum"]) and hi == float.fromhex(case["maximum"])


def test_two_opt_preparation_golden():
    # checklist 2-opt from a given tour with every city queued in tour order
    # (the initial local search of preparation), with evaluation limits
    for case in load("two_opt.json"):
        tour, cost, evals = O.two_opt(np.array(case["xy"]), np.array(case["ls"]), case["tour"], case["tour"],
                                      case["limit"])
        assert tour.tolist() == case["result"]
        assert cost == case["cost"] and evals == case["evaluations"]


def test_prepare_rows_brute_force_and_nn_tour():
    rng = np.random.default_rng(8)
    for n, grid in ((9, 10000), (40, 6), (150, 10000)):
        xy = rng.integers(0, grid + 1, size=(n, 2))
        M = O.distance_matrix(xy)
        p = O.prepare(xy, O.default_settings())
        s = p["settings"]
        for a in range(n):
            order = sorted((b for b in range(n) if b != a), key=lambda b: (M[a, b], b))
            assert p["primary"][a].tolist() == order[: s.primary_width]
            assert p["backup"][a].tolist() == order[s.primary_width: s.primary_width + s.backup_width]
            assert p["ls"][a].tolist() == order[: s.ls_width]
            k = min(8, n - 1)
            assert p["local_scale"][a] == sum(M[a, b] / k for b in order[:k])
        tour = p["initial_tour"]
        assert sorted(tour.tolist()) == list(range(n))
        assert p["initial_cost"] == int(M[tour, np.roll(tour, -1)].sum())
        assert p["initial_cost"] <= p["cheap_cost"]


def test_pheromone_golden_replayed_by_oracle_colony_rule():
    # The oracle's update (evaporate, clamp, deposit on parent edges, clamp,
    # products) over several steps equals the reference SparsePheromone.
    for case in load("pheromone.json"):
        xy = np.array(case["xy"])
        M = O.distance_matrix(xy)
        primary = np.array(case["primary"])
        n, pw = primary.shape
        trails = np.full(n * pw, float.fromhex(case["initial"]))
        rho = case["retention"]
        for step, parent in enumerate(case["parents"]):
            lo, hi = float.fromhex(case["minima"][step]), float.fromhex(case["maxima"][step])
            inc = float.fromhex(case["increments"][step])
            pos = np.argsort(parent)
            for a in range(n):
                succ, pred = parent[(pos[a] + 1) % n], parent[(pos[a] - 1) % n]
                for j in range(pw):
                    v = max(lo, trails[a * pw + j] * rho)
                    if primary[a, j] in (succ, pred):
                        v = min(hi, v + inc)
                    trails[a * pw + j] = v
            want = np.array([float.fromhex(v) for v in case["trails"][step]])
            np.testing.assert_array_equal(trails, want)
            eta = np.array([1.0 / M[a, primary[a, j]] if M[a, primary[a, j]] else 1.0
                            for a in range(n) for j in range(pw)])
            np.testing.assert_array_equal(trails * eta, np.array([float.fromhex(v) for v in case["products"][step]]))


def test_run_faco_trace_pheromone_matches_reference_rule():
    """Oracle colony run: its per-batch trails equal the SparsePheromone rule
    replayed on its own parents and epoch costs (bounds from the epoch cost,
    deposit 1/parent cost), and costs never increase."""
    rng = np.random.default_rng(12)
    n = 40
    xy = rng.integers(0, 10001, size=(n, 2)).astype(np.int32)
    s = O.default_settings(ants=8, primary_width=6, backup_width=4, retention=0.75, epoch_source_probability=0.3)
    run = O.run_faco(xy[None], [55], [3], s, 12, 8, threads=1, trace=True)
    prep = O.prepare(xy, s)
    pw = prep["settings"].primary_width
    M = O.distance_matrix(xy)
    trails = np.full(n * pw, O.trail_limits(pw, s.p_best, s.retention, float(prep["initial_cost"]))[1])
    last_global = prep["initial_cost"]
    for b in range(12):
        epoch, glob, parent_cost = run["trace_costs"][b]
        assert glob <= last_global and epoch >= glob
        last_global = glob
        parent = run["trace_parent"][b]
        assert int(M[parent, np.roll(parent, -1)].sum()) == parent_cost
        lo, hi = O.trail_limits(pw, s.p_best, s.retention, float(epoch))
        pos = np.argsort(parent)
        for a in range(n):
            succ, pred = parent[(pos[a] + 1) % n], parent[(pos[a] - 1) % n]
            for j in range(pw):
                v = max(lo, trails[a * pw + j] * s.retention)
                if prep["primary"][a, j] in (succ, pred):
                    v = min(hi, v + 1.0 / parent_cost)
                trails[a * pw + j] = v
        np.testing.assert_array_equal(run["trace_trails"][b], trails)
    assert run["costs"][0] == last_global


# ---------------------------------------------------------------------------
# live reference checker (oracle/_ref), where it is built

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checker oracle/_ref not built here")


@needs_ref
@pytest.mark.parametrize("n", [3, 5, 9, 33, 120])
def test_construct_ls_against_reference_core(n):
    rng = np.random.default_rng(1000 + n)
    for metric in ("uniform", "grid", "repeated"):
        if metric == "uniform":
            xy = rng.integers(0, 10001, size=(n, 2))
        elif metric == "grid":
            xy = rng.integers(0, 9, size=(n, 2))
        else:
            xy = rng.integers(0, 10001, size=(n, 2))
            xy[2::2] = xy[1]
        M = O.distance_matrix(xy)
        width = min(20, n - 1)
        ls = np.array([sorted((b for b in range(n) if b != a), key=lambda b: (M[a, b], b))[:width]
                       for a in range(n)])
        for cap in (0, 1, 7, 127, 2**64 - 1):
            for mne in (2, 8, 16):
                parent, visit = rng.permutation(n), rng.permutation(n)
                a = O.construct_ls(xy, ls, parent, visit, mne, cap)
                b = O.ref_construct_ls(M.astype(np.float64), ls, parent, visit, mne, cap)
                for key in ("construction_tour", "tour", "positions", "checklist"):
                    assert np.array_equal(a[key], b[key]), key
                assert a["counters"] == b["counters"]
                assert a["final_cost"] == b["final_cost"] and a["construction_cost"] == b["construction_cost"]


@needs_ref
def test_select_next_and_limits_against_reference_core():
    rng = np.random.default_rng(4)
    for k in range(300):
        n = int(rng.integers(4, 25))
        xy = rng.integers(0, 50, size=(n, 2))
        M = O.distance_matrix(xy)
        cur = int(rng.integers(0, n))
        order = sorted((b for b in range(n) if b != cur), key=lambda b: (M[cur, b], b))
        pw = min(int(rng.integers(2, 8)), n - 1)
        bw = min(int(rng.integers(0, 5)), n - 1 - pw)
        visited = (rng.random(n) < rng.random()).astype(np.uint8)
        visited[cur] = 1
        products = rng.random(pw) * (k % 5 != 0)
        u = float(rng.random())
        assert O.select_next(xy, order[:pw], order[pw:pw + bw], products, visited, cur, u) == \
            O.ref_select_next(M.astype(np.float64), order[:pw], order[pw:pw + bw], products, visited, cur, u)
    for width in (2, 16, 20):
        for cost in (10.0, 232511.0):
            assert O.trail_limits(width, 0.1, 0.5, cost) == O.ref_trail_limits(width, 0.1, 0.5, cost)
