"""The oracle's control layer (restart and perturbation control) against the
reference's control code.

- golden vectors written from the reference's control_cpu.cpp /
  control_ops.hpp / baseline_policy.hpp / factorial_policy.hpp
  (tests/golden/control.json, oracle/_ref) and the live checker where built;
- the hand-derived cases of the reference's tests/cpp/test_control_cpu.cpp
  and tests/cpp/test_baseline_policy.cpp;
- composition invariants of the controlled run (per-batch order of
  cuda/batch_engine.cu:880-1060), which has no CPU counterpart in the
  reference.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "control.json").read_text())
POLICY_FIELDS = ("kind", "mne_level", "max_mne_level", "region", "restart_mode", "restart_period",
                 "restart_probability", "stagnation_step", "restart_stagnation", "restart_cooldown")


def policy(values):
    b = O.Baseline()
    for name, v in zip(POLICY_FIELDS, values):
        setattr(b, name, v)
    return b


def policy_of(**kw):
    b = O.Baseline()
    for name, v in kw.items():
        setattr(b, name, v)
    return b


def fb_of(stagnant=0, epoch=0, ret=0.0, work=0.0, restarts=0):
    return O.Feedback(ret, work, stagnant, epoch, restarts)


# ---------------------------------------------------------------------------
# golden vectors

def test_identity_and_canonical_order_golden():
    for case in GOLDEN["identity"]:
        a, b = np.array(case["a"], np.uint32), np.array(case["b"], np.uint32)
        assert [str(v) for v in O.tour_ident(a)] == case["id"]
        assert O.canonical_compare(a, b) == case["compare"]
        assert O.canonical_compare(a, np.roll(a, 3)[::-1].copy()) == case["compare_rotated"] == 0


def test_sample_nodes_and_regions_golden():
    for case in GOLDEN["samples"]:
        assert O.sample_nodes(case["n"], int(case["key"])).tolist() == case["nodes"]
    for case in GOLDEN["regions"]:
        tour, primary = np.array(case["tour"], np.uint32), np.array(case["primary"], np.uint32)
        for m in range(2):
            for r in range(4):
                got = O.build_region(tour, primary, int(case["key"]), case["batch"], m, r)
                assert got.tolist() == case["regions"][m][r]


def test_archive_update_golden():
    for case in GOLDEN["archive"]:
        n = len(case["global"])
        got = O.archive_update(np.array(case["xy"]), np.array(case["global"], np.uint32),
                               np.array(case["entries"], np.uint32).reshape(-1, n),
                               np.array(case["completed"], np.uint32).reshape(-1, n))
        assert got.tolist() == case["chosen"]


def test_action_snapshot_golden():
    for case in GOLDEN["snapshot"]:
        n = len(case["active"])
        f = case["feedback"]
        out = O.action_snapshot(np.array(case["xy"]), np.array(case["primary"], np.uint32),
                                np.array(case["local_scale"]), 1e-12, np.array(case["samples"], np.uint32),
                                np.array(case["active"], np.uint32), np.array(case["archive"], np.uint32),
                                np.array([float.fromhex(v) for v in case["trails"]]).reshape(n, -1),
                                float.fromhex(case["default"]), float.fromhex(case["minimum"]),
                                float.fromhex(case["maximum"]), O.Feedback(f[0], f[1], f[2], f[3], f[4]),
                                float.fromhex(case["elapsed"]), int(case["key"]), case["batch"], case["mask"])
        assert out["features"].view(np.uint32).reshape(-1).tolist() == case["features_bits"]
        assert out["regions"].tolist() == case["regions"]
        assert out["alternative"] == case["alternative"] and out["mask"] == case["legal_mask"]


def test_policies_and_feedback_golden():
    for case in GOLDEN["policy"]:
        b = policy(case["policy"])
        fb = fb_of(*case["feedback"])
        u = float.fromhex(case["u"])
        assert O.baseline_action(b, fb, case["batch"], u, case["mask"]) == case["action"]
        assert [O.factorial_mask(v, b, fb, case["batch"], u, case["legal"]) for v in range(4)] == case["factorial"]
    for case in GOLDEN["feedback"]:
        r, w, st, ep, rs = case["before"]
        fb = O.Feedback(r, w, st, ep, rs)
        improved, ret, work, ants = case["args"]
        O.update_feedback(fb, improved, float.fromhex(ret), float.fromhex(work), ants)
        if case["restart"]:
            O.restart_feedback(fb)
        got = [int(np.float32(fb.return_rate).view(np.uint32)), int(np.float32(fb.ls_work).view(np.uint32)),
               fb.stagnant_batches, fb.epoch_batches, fb.restarts]
        assert got == case["after_bits"]


# ---------------------------------------------------------------------------
# hand-derived cases of the reference's own tests

def test_baseline_hand_cases():
    # tests/cpp/test_baseline_policy.cpp:20-73
    for region in range(4):
        for level in range(4):
            p = policy_of(region=region, mne_level=level, max_mne_level=level)
            f = fb_of(stagnant=2**64 - 1, epoch=93)
            keep = region * 4 + level
            assert O.baseline_action(p, f, 23, 0, 0xFFFFFFFF) == keep
            p.restart_mode, p.restart_probability = 2, 1.0
            assert O.baseline_action(p, f, 23, 0.999999, 0xFFFFFFFF) == keep + 16
            assert O.baseline_action(p, f, 23, 0, 0xFFFF) == keep
    periodic = policy_of(restart_mode=1, restart_period=3)
    assert [O.baseline_action(periodic, fb_of(), b, 0, 0xFFFFFFFF) for b in (0, 1, 2, 4, 5)] == [0] * 5
    assert [O.baseline_action(periodic, fb_of(), b, 0, 0xFFFFFFFF) for b in (3, 6, 9)] == [16] * 3
    bern = policy_of(restart_mode=2, restart_probability=0.5)
    assert O.baseline_action(bern, fb_of(), 0, 0.499999, 0xFFFFFFFF) == 16
    assert O.baseline_action(bern, fb_of(), 0, 0.5, 0xFFFFFFFF) == 0
    rule = policy_of(kind=1, region=2, mne_level=1, max_mne_level=3, stagnation_step=4, restart_stagnation=9,
                     restart_cooldown=3)
    for stagnant, epoch, expected in ((0, 0, 9), (3, 3, 9), (4, 4, 10), (7, 7, 10), (8, 8, 11), (9, 2, 11),
                                      (9, 3, 27), (2**64 - 1, 2**64 - 1, 27)):
        assert O.baseline_action(rule, fb_of(stagnant, epoch), 0, 0, 0xFFFFFFFF) == expected
    f = fb_of(9, 9)
    O.restart_feedback(f)
    assert (f.stagnant_batches, f.epoch_batches) == (9, 0)
    assert O.baseline_action(rule, f, 100, 0, 0xFFFFFFFF) == 11
    O.update_feedback(f, True, 0, 0, 1)
    assert O.baseline_action(rule, f, 101, 0, 0xFFFFFFFF) == 9


def test_feedback_hand_cases():
    # tests/cpp/test_control_cpu.cpp:113-124
    f = fb_of()
    O.update_feedback(f, False, 1, 0.5, 32)
    assert abs(f.return_rate - 1 / 16) < 1e-7 and abs(f.ls_work - 1 / 32) < 1e-7
    assert (f.stagnant_batches, f.epoch_batches) == (1, 1)
    O.update_feedback(f, True, 0, 0, 0)
    assert (f.stagnant_batches, f.epoch_batches) == (1, 1)
    O.restart_feedback(f)
    assert (f.return_rate, f.ls_work, f.epoch_batches, f.stagnant_batches, f.restarts) == (0, 0, 0, 1, 1)
    f.stagnant_batches = f.epoch_batches = 2**64 - 1
    O.update_feedback(f, False, 0, 0, 1)
    assert f.stagnant_batches == f.epoch_batches == 2**64 - 1


def test_factorial_mask_closed_forms():
    # factorial_policy.hpp:32-45: M11 keeps the legal mask, M00 is the
    # baseline's action, M01 keeps the baseline's region/level in both modes,
    # M10 keeps the baseline's mode and falls back to holding.
    rng = np.random.default_rng(8)
    for _ in range(300):
        b = policy_of(region=int(rng.integers(4)), mne_level=int(rng.integers(4)), restart_mode=2,
                      restart_probability=float(rng.random()))
        b.max_mne_level = b.mne_level
        f, batch, u = fb_of(int(rng.integers(10)), int(rng.integers(10))), int(rng.integers(9)), float(rng.random())
        legal = int(rng.integers(1, 2**32)) | 1
        assert O.factorial_mask(3, b, f, batch, u, legal) == legal
        assert O.factorial_mask(0, b, f, batch, u, legal) == 1 << O.baseline_action(b, f, batch, u, legal)
        requested = O.baseline_action(b, f, batch, u, 0xFFFFFFFF)
        keep = requested % 16
        assert O.factorial_mask(2, b, f, batch, u, legal) == legal & ((1 << keep) | (1 << (keep + 16)))
        half = legal & (0xFFFF0000 if requested >= 16 else 0xFFFF)
        assert O.factorial_mask(1, b, f, batch, u, legal) == (half or legal & 0xFFFF)


def test_identity_rotation_reversal_and_regions():
    # tests/cpp/test_control_cpu.cpp:22-77
    order = np.arange(17, dtype=np.uint32)
    base = O.tour_ident(order)
    for shift in range(17):
        for rev in (False, True):
            t = np.roll(order, -shift)
            t = t[::-1].copy() if rev else t
            assert O.tour_ident(t) == base and O.canonical_compare(t, order) == 0
    changed = order.copy()
    changed[[5, 8]] = changed[[8, 5]]
    assert O.tour_ident(changed) != base and O.canonical_compare(changed, order) != 0
    rng = np.random.default_rng(49178)
    for n in (3, 7, 16, 17, 64, 100):
        tour = np.arange(n, dtype=np.uint32)
        primary = np.array([[a - 1 if a % 2 else (a + 1 if a + 1 < n else a - 1)] for a in range(n)], np.uint32)
        padded = np.full((n, 4), n, dtype=np.uint32)
        padded[:, 0] = primary[:, 0]
        for repeat in range(10):
            key = int(rng.integers(0, 2**63))
            samples = O.sample_nodes(n, key)
            assert len(set(samples.tolist())) == min(n, 64) and samples.max() < n
            for used in range(min(n, 64)):
                remaining = [v for v in range(n) if v not in samples[:used]]
                for rank in range(0, len(remaining), max(1, len(remaining) // 7)):
                    assert O.remaining_node(rank, samples[:used]) == remaining[rank]
            for mode in range(2):
                for region in range(4):
                    nodes = O.build_region(tour, primary, key, repeat, mode, region)
                    r = min(n, 16)
                    assert len(set(nodes.tolist())) == r and nodes.max() < n
                    assert (O.build_region(tour, padded, key, repeat, mode, region) == nodes).all()
                    if region < 2:
                        assert all(nodes[i] == (nodes[i - 1] + 1) % n for i in range(1, r))


# ---------------------------------------------------------------------------
# live reference checker

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checker oracle/_ref not built here")


@needs_ref
def test_control_primitives_against_reference_core():
    rng = np.random.default_rng(11)
    for n in (3, 5, 17, 100, 1000):
        for k in range(10):
            t = rng.permutation(n).astype(np.uint32)
            u = rng.permutation(n).astype(np.uint32)
            assert O.tour_ident(t) == O.ref_fingerprint(t)
            assert O.canonical_compare(t, u) == O.ref_canonical_compare(t, u)
            key = int(rng.integers(0, 2**63))
            assert (O.sample_nodes(n, key) == O.ref_sample_nodes(n, key)).all()
            primary = np.array([rng.permutation([j for j in range(n) if j != a])[:min(4, n - 1)]
                                for a in range(n)], dtype=np.uint32) if n <= 100 else \
                np.stack([(np.arange(n) + j + 1) % n for j in range(4)], axis=1).astype(np.uint32)
            for mode in range(2):
                for region in range(4):
                    assert (O.build_region(t, primary, key, k, mode, region) ==
                            O.ref_build_region(t, primary, key, k, mode, region)).all()


@needs_ref
def test_archive_and_snapshot_against_reference_core():
    rng = np.random.default_rng(21)
    for n in (6, 12, 30):
        for rep in range(10):
            xy = rng.integers(0, 200 if rep % 2 else 10001, size=(n, 2))
            M = O.distance_matrix(xy)
            prep = O.prepare(xy, O.default_settings(primary_width=min(4, n - 1), backup_width=0,
                                                    ls_width=min(6, n - 1)))
            g = prep["initial_tour"].astype(np.uint32)
            pool = [g]
            for _ in range(10):
                t = g.copy()
                i, j = sorted(rng.choice(n, 2, replace=False))
                t[i:j + 1] = t[i:j + 1][::-1]
                pool.append(np.roll(t, int(rng.integers(n))).astype(np.uint32))
            pool += [np.roll(g, 2).astype(np.uint32), g[::-1].copy()]
            costs = [int(M[t, np.roll(t, -1)].sum()) for t in pool]
            gi = int(np.argmin(costs))
            rest = np.array([t for k, t in enumerate(pool) if k != gi])
            k = int(rng.integers(0, 4))
            a = O.archive_update(xy, pool[gi], rest[:k], rest[k:])
            b = O.ref_archive_update(M, pool[gi], rest[:k], rest[k:])
            assert a.tolist() == b.tolist()
            order = np.argsort(costs, kind="stable")
            active = pool[int(rng.integers(len(pool)))]
            key, batch = int(rng.integers(0, 2**63)), int(rng.integers(0, 500))
            samples = O.sample_nodes(n, key)
            fb = O.Feedback(float(rng.random()), float(rng.random()), int(rng.integers(0, 40)), 2, 0)
            deps = np.array([rng.permutation(n)], dtype=np.uint32)
            lo, hi = sorted(float(v) for v in rng.random(2) * 0.01)
            r = O.ref_action_snapshot(M, prep["primary"], prep["local_scale"], 1e-12, samples, active,
                                      np.array([pool[i] for i in order]), 0.005, 0.7, 1e-4, deps, [0.003], lo, hi,
                                      fb, 0.4, key, batch)
            o = O.action_snapshot(xy, prep["primary"], prep["local_scale"], 1e-12, samples, active, r["archive"],
                                  r["trails"], r["default"], lo, hi, fb, 0.4, key, batch)
            assert (o["features"].view(np.uint32) == r["features"].view(np.uint32)).all()
            assert (o["regions"] == r["regions"]).all() and o["alternative"] == r["alternative"]


# ---------------------------------------------------------------------------
# controlled run: composition invariants

def instances(n, count, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 10001, size=(count, n, 2)).astype(np.int32)


@pytest.fixture(scope="module")
def panel():
    return instances(40, 2, seed=3), np.array([101, 202], dtype=np.uint64)


RESTARTY = O.make_program([0, 0, 2], [4, 1, 0])       # mode + stagnation: prefers restarts
HOLDY = O.make_program([0, 1, 3], [6, 0, 0], [0.5])   # gap - 0.5


def test_controller_kind0_and_factorial_endpoints_match(panel):
    xy, keys = panel
    seeds = np.array([1, 2], dtype=np.uint64)
    s = O.default_settings(ants=8)
    limit = 8 * 12
    prog = O.run_controlled(xy, keys, seeds, s, limit, "program", programs=[RESTARTY], trace=True)
    c = O.Controller()
    c.kind, c.trees[0] = 0, RESTARTY
    ctrl = O.run_controlled(xy, keys, seeds, s, limit, "controller", controllers=[c], trace=True)
    assert (prog["tours"] == ctrl["tours"]).all() and (prog["trace_actions"] == ctrl["trace_actions"]).all()
    m11 = O.run_controlled(xy, keys, seeds, s, limit, "factorial", programs=[RESTARTY],
                           baseline=policy_of(restart_mode=2, restart_probability=0.3), variant="M11")
    assert (m11["tours"] == prog["tours"]).all()
    base = policy_of(region=1, mne_level=2, max_mne_level=2, restart_mode=2, restart_probability=0.3)
    b = O.run_controlled(xy, keys, seeds, s, limit, "baseline", baseline=base, trace=True)
    m00 = O.run_controlled(xy, keys, seeds, s, limit, "factorial", programs=[HOLDY], baseline=base, variant="M00",
                           trace=True)
    assert (b["tours"] == m00["tours"]).all() and (b["trace_actions"] == m00["trace_actions"]).all()
    assert set(np.unique(b["trace_actions"])) <= {6, 22}


def test_population_duplicates_and_trace_closed_forms(panel):
    xy, keys = panel
    s = O.default_settings(ants=8)
    limit = 8 * 10
    # colonies = 2 instances x 3 individuals; individuals 0 and 2 share a program
    seeds = np.array([5, 6] * 3, dtype=np.uint64)
    out = O.run_controlled(xy, keys, seeds, s, limit, "program", programs=[RESTARTY, HOLDY, RESTARTY], trace=True)
    assert (out["tours"][0:2] == out["tours"][4:6]).all() and (out["costs"][0:2] == out["costs"][4:6]).all()
    f, actions, masks = out["trace_features"], out["trace_actions"], out["trace_masks"]
    batches = f.shape[0]
    for b in range(batches):
        assert np.all(f[b, 0] == np.float32(min(1.0, b * 8 / limit)))
        assert np.all(f[b, 4] == (np.arange(32) >= 16).astype(np.float32))
        assert np.all(f[b, 5] == ((np.arange(32) % 4 + 1) * 0.25).astype(np.float32))
        for c in range(6):
            assert (masks[b, c] >> actions[b, c]) & 1
            assert (masks[b, c] == 0xFFFFFFFF) == (out["trace_alternatives"][b, c] < 4)
    # batch 0: the archive holds only the initial tour, which is the parent
    assert np.all(out["trace_alternatives"][0] == 4) and np.all(actions[0] < 16)
    restarts = (actions >= 16).sum(axis=0)
    assert [st["restarts"] for st in out["states"]] == restarts.tolist()
    initial = [O.prepare(xy[g], s)["initial_cost"] for g in range(2)]
    for c in range(6):
        # a restart zeroes epoch_batches before the batch's own increment
        last = np.nonzero(actions[:, c] >= 16)[0]
        assert out["states"][c]["epoch_batches"] == (batches - last[-1] if last.size else batches)
        # stagnation counts batches since the global cost last fell
        g = np.concatenate([[initial[c % 2]], out["trace_costs"][:, c, 1]])
        improved = np.nonzero(g[1:] < g[:-1])[0]
        assert out["states"][c]["stagnant_batches"] == (batches - 1 - improved[-1] if improved.size else batches)


def test_hold_only_mask_never_restarts(panel):
    xy, keys = panel
    s = O.default_settings(ants=8)
    out = O.run_controlled(xy, keys, np.array([1, 2], dtype=np.uint64), s, 8 * 10, "program",
                           programs=[RESTARTY], experiment_mask=0xFFFF, trace=True)
    assert np.all(out["trace_actions"] < 16) and all(st["restarts"] == 0 for st in out["states"])
    with pytest.raises(ValueError):
        O.run_controlled(xy, keys, np.array([1, 2], dtype=np.uint64), s, 8 * 10, "program",
                         programs=[RESTARTY], experiment_mask=0xFFFF0000)


# ---------------------------------------------------------------------------
# stopping early and paired forks (the oracle side of state snapshots,
# reference cuda/batch_engine.cu:839-845 and 883-923)

def _same_prefix(a, b, batches):
    for name in ("features", "scores", "masks", "actions", "alternatives", "regions", "costs"):
        x, y = a["trace_" + name][:batches], b["trace_" + name][:batches]
        assert (x.view(np.uint8) == y.view(np.uint8)).all(), name


def test_stop_batch_is_the_prefix_of_the_full_run(panel):
    xy, keys = panel
    seeds = np.array([4, 9], dtype=np.uint64)
    s = O.default_settings(ants=8)
    full = O.run_controlled(xy, keys, seeds, s, 8 * 9, "program", programs=[RESTARTY], trace=True)
    for stop in (1, 5):
        part = O.run_controlled(xy, keys, seeds, s, 8 * 9, "program", programs=[RESTARTY], trace=True,
                                stop_batch=stop)
        assert part["trace_actions"].shape[0] == stop
        _same_prefix(part, full, stop)
        assert part["costs"].tolist() == full["trace_costs"][stop - 1, :, 1].tolist()
    # the progress feature (feature 0) runs over the whole budget, not the stop
    assert abs(float(part["trace_features"][4, 0, 0, 0]) - 4 / 9) < 1e-6


@pytest.mark.parametrize("region,level", [(0, 0), (2, 3)])
def test_fork_forces_one_decision_and_rekeys_the_ants(panel, region, level):
    xy, keys = panel
    seeds = np.array([4, 9], dtype=np.uint64)
    s = O.default_settings(ants=8)
    base = policy_of(kind=0, mne_level=1, max_mne_level=1, region=1)
    full = O.run_controlled(xy, keys, seeds, s, 8 * 9, "baseline", baseline=base, trace=True)
    cut, action = 4, region * 4 + level
    fork = O.run_controlled(xy, keys, seeds, s, 8 * 9, "baseline", baseline=base, trace=True, stop_batch=7,
                            fork=(cut, 53, action))
    _same_prefix(fork, full, cut)
    # the forced batch's decision inputs are those of the unforked run: the
    # regions and features are drawn before the keys change
    for name in ("features", "masks", "alternatives", "regions"):
        assert (fork["trace_" + name][cut].view(np.uint8) == full["trace_" + name][cut].view(np.uint8)).all(), name
    assert fork["trace_actions"][cut].tolist() == [action] * 2
    assert full["trace_actions"][cut].tolist() == [1 * 4 + 1] * 2
    assert fork["trace_actions"][cut + 1:].tolist() == full["trace_actions"][cut + 1:7].tolist()
    # later regions come from the new keys
    assert not (fork["trace_regions"][cut + 1:7] == full["trace_regions"][cut + 1:7]).all()
    again = O.run_controlled(xy, keys, seeds, s, 8 * 9, "baseline", baseline=base, stop_batch=7, fork=(cut, 53, action))
    assert again["tours"].tolist() == fork["tours"].tolist() and again["work"].tolist() == fork["work"].tolist()
    other = O.run_controlled(xy, keys, seeds, s, 8 * 9, "baseline", baseline=base, trace=True, stop_batch=7,
                             fork=(cut, 54, action))
    assert (other["trace_features"][cut].view(np.uint32) == fork["trace_features"][cut].view(np.uint32)).all()
    assert not (other["trace_regions"][cut + 1:] == fork["trace_regions"][cut + 1:]).all()


def test_alternate_stream_is_unrelated_and_restored():
    xy = instances(60, 1, seed=8)
    s = O.default_settings(ants=8)
    seeds = np.array([3, 4], dtype=np.uint64)
    before = O.run_faco(xy, [9], seeds, s, 5, 8)
    with O.alternate_stream():
        other = O.run_faco(xy, [9], seeds, s, 5, 8)
        again = O.run_faco(xy, [9], seeds, s, 5, 8)
        assert O.uniform53(9, 0, 0, 0) != 0.0
    after = O.run_faco(xy, [9], seeds, s, 5, 8)
    assert before["tours"].tolist() == after["tours"].tolist()
    assert other["tours"].tolist() == again["tours"].tolist()
    assert other["work"].tolist() != before["work"].tolist()


def test_progress_override_reproduces_the_count_feature():
    """The explicit per-batch progress feature (time-budget runs) equals the
    default run when it carries the default's values, completed FE over the
    evaluation limit, and changes the run when it does not."""
    rng = np.random.default_rng(61)
    n, ants, batches = 45, 8, 12
    xy = rng.integers(0, 10001, size=(2, n, 2)).astype(np.int32)
    s = O.default_settings(ants=ants)
    program = O.make_program([0, 0, 2], [0, 1, 0], feature_spec_id=1)   # progress + stagnation
    seeds = np.array([3, 4], dtype=np.uint64)
    limit = ants * batches
    base = O.run_controlled(xy, [11, 12], seeds, s, limit, "program", programs=[program], trace=True)
    same = O.run_controlled(xy, [11, 12], seeds, s, limit, "program", programs=[program], trace=True,
                            progress=[b * ants / limit for b in range(batches)])
    np.testing.assert_array_equal(base["tours"], same["tours"])
    np.testing.assert_array_equal(base["trace_features"], same["trace_features"])
    other = O.run_controlled(xy, [11, 12], seeds, s, limit, "program", programs=[program], trace=True,
                             progress=[1.0] * batches)
    assert (other["trace_features"][:, 0] == 1.0).all()
    assert not np.array_equal(other["trace_actions"], base["trace_actions"]) or \
        not np.array_equal(other["trace_features"], base["trace_features"])
