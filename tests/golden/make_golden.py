"""Write golden vectors from the reference's own CPU core (oracle/_ref).

Run in a checkout that has the reference sources (the recipe in
oracle/Makefile builds oracle/_ref/libgpfaco_ref.so from them):

    make -C oracle ref && python tests/golden/make_golden.py

Inputs are generated here from fixed seeds; expected outputs come from the
reference's FocusedConstruction + checklist_two_opt, select_next,
candidate_trail_limits, SparsePheromone, score_cpu and the control layer
(fingerprint, canonical order, sample_nodes, build_region, ReferenceArchive,
make_action_snapshot, update/restart_feedback, baseline_action,
factorial_mask), called with the integer EUC_2D distances this engine uses.
The fixtures are data only.
"""
from __future__ import annotations

import json
import struct
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from gplsaco_amd.programs import random_program  # noqa: E402

OUT = Path(__file__).resolve().parent
U64 = (1 << 64) - 1


def rows(M, width):
    n = M.shape[0]
    out = np.zeros((n, width), dtype=np.uint32)
    for a in range(n):
        out[a] = sorted((b for b in range(n) if b != a), key=lambda b: (M[a, b], b))[:width]
    return out


def points(rng, n, metric):
    if metric == "uniform":
        return rng.integers(0, 10001, size=(n, 2))
    if metric == "grid":
        return rng.integers(0, 13, size=(n, 2))
    xy = rng.integers(0, 10001, size=(n, 2))
    xy[3::3] = xy[0]
    return xy


def construct_ls_cases():
    rng = np.random.default_rng(20240607)
    cases = []
    for n in (3, 4, 5, 7, 8, 17, 40):
        for metric in ("uniform", "grid", "repeated"):
            xy = points(rng, n, metric)
            M = O.distance_matrix(xy)
            ls = rows(M, min(20, n - 1))
            for cap, mne in ((0, 2), (1, 4), (7, 8), (127, 16), (U64, 8)):
                parent = rng.permutation(n)
                visit = rng.permutation(n)
                r = O.ref_construct_ls(M.astype(np.float64), ls, parent, visit, mne, cap)
                cases.append({
                    "xy": xy.tolist(), "ls": ls.tolist(), "parent": parent.tolist(), "visit": visit.tolist(),
                    "mne_target": mne, "limit": cap,
                    "construction_tour": r["construction_tour"].tolist(), "tour": r["tour"].tolist(),
                    "positions": r["positions"].tolist(), "checklist": r["checklist"].tolist(),
                    "counters": r["counters"], "construction_cost": r["construction_cost"],
                    "final_cost": r["final_cost"],
                })
    return cases


def select_next_cases():
    rng = np.random.default_rng(77)
    cases = []
    for k in range(120):
        n = int(rng.integers(5, 30))
        xy = points(rng, n, "grid" if k % 3 == 0 else "uniform")
        M = O.distance_matrix(xy)
        current = int(rng.integers(0, n))
        order = sorted((b for b in range(n) if b != current), key=lambda b: (M[current, b], b))
        pw = min(int(rng.integers(2, 8)), n - 1)
        bw = min(int(rng.integers(0, 6)), n - 1 - pw)
        primary, backup = order[:pw], order[pw:pw + bw]
        visited = (rng.random(n) < [0.3, 0.7, 0.95][k % 3]).astype(np.uint8)
        visited[current] = 1
        products = rng.random(pw) * (0.0 if k % 7 == 0 else 1.0)
        u = float(rng.random())
        chosen, stage = O.ref_select_next(M.astype(np.float64), primary, backup, products, visited, current, u)
        cases.append({"xy": xy.tolist(), "current": current, "primary": primary, "backup": backup,
                      "products": products.tolist(), "visited": visited.tolist(), "uniform": u,
                      "chosen": chosen, "stage": stage})
    return cases


def trail_limit_cases():
    cases = []
    for width in (2, 3, 16, 20, 64):
        for p_best in (0.05, 0.1, 0.5):
            for rho in (0.0, 0.25, 0.5, 0.9):
                for cost in (1.0, 77.0, 232511.0, 1e7):
                    lo, hi = O.ref_trail_limits(width, p_best, rho, cost)
                    cases.append({"width": width, "p_best": p_best, "retention": rho, "cost": cost,
                                  "minimum": lo.hex(), "maximum": hi.hex()})
    return cases


def pheromone_cases():
    rng = np.random.default_rng(91)
    cases = []
    for n, pw, rho in ((9, 2, 0.5), (16, 4, 0.25), (30, 16, 0.9)):
        xy = points(rng, n, "uniform")
        M = O.distance_matrix(xy)
        primary = rows(M, pw)
        steps = 5
        parents = np.stack([rng.permutation(n) for _ in range(steps)])
        costs = [float(M[t, np.roll(t, -1)].sum()) for t in parents]
        limits = [O.ref_trail_limits(pw, 0.1, rho, c) for c in costs]
        initial = limits[0][1]
        minima = [lim[0] for lim in limits]
        maxima = [lim[1] for lim in limits]
        increments = [1.0 / c for c in costs]
        trails, defaults, products = O.ref_pheromone_run(M.astype(np.float64), primary, initial, parents, rho,
                                                         minima, maxima, increments)
        cases.append({"xy": xy.tolist(), "primary": primary.tolist(), "retention": rho, "initial": initial.hex(),
                      "parents": parents.tolist(), "minima": [m.hex() for m in minima],
                      "maxima": [m.hex() for m in maxima], "increments": [i.hex() for i in increments],
                      "trails": [[v.hex() for v in row] for row in trails],
                      "defaults": [d.hex() for d in defaults],
                      "products": [[v.hex() for v in row] for row in products]})
    return cases


def two_opt_cases():
    rng = np.random.default_rng(5150)
    cases = []
    for n in (6, 12, 30, 60):
        for metric in ("uniform", "grid"):
            xy = points(rng, n, metric)
            M = O.distance_matrix(xy)
            ls = rows(M, min(20, n - 1))
            tour = rng.permutation(n)
            for limit in (3, 50, 100000):
                out, cost, evals = O.ref_two_opt(M.astype(np.float64), ls, tour, tour, limit)
                cases.append({"xy": xy.tolist(), "ls": ls.tolist(), "tour": tour.tolist(), "limit": limit,
                              "result": out.tolist(), "cost": cost, "evaluations": evals})
    return cases


def score_cases():
    rng = np.random.default_rng(6677)
    features = rng.random((12, 5, 32), dtype=np.float32)
    masks = rng.integers(1, 0xFFFFFFFF, size=5, dtype=np.uint32)
    masks[:3] = [1, 1 << 31, 0xFFFFFFFF]
    cases = []
    for _ in range(60):
        program = random_program(rng, depth=5)
        constants = [struct.unpack("<f", struct.pack("<I", b))[0] for b in program["constant_bits"]]
        scores, actions = O.ref_score_cpu(program["opcode"], program["operand"], constants, features, masks)
        cases.append({"program": program, "scores_bits": scores.view(np.uint32).tolist(),
                      "actions": actions.tolist()})
    return {"features_bits": features.view(np.uint32).tolist(), "masks": masks.tolist(), "cases": cases}


def tour_variants(rng, base, count):
    n = base.size
    out = []
    for _ in range(count):
        t = base.copy()
        i, j = sorted(rng.choice(n, 2, replace=False))
        t[i:j + 1] = t[i:j + 1][::-1]
        if rng.random() < 0.5:
            t = np.roll(t, int(rng.integers(n)))
        if rng.random() < 0.3:
            t = t[::-1].copy()
        out.append(t.astype(np.uint32))
    return out


def random_policy(rng):
    b = O.Baseline()
    b.kind = int(rng.integers(0, 2))
    b.region, b.mne_level = int(rng.integers(0, 4)), int(rng.integers(0, 4))
    if b.kind == 0:
        b.max_mne_level = b.mne_level
        b.restart_mode = int(rng.integers(0, 3))
        if b.restart_mode == 1:
            b.restart_period = int(rng.integers(1, 7))
        if b.restart_mode == 2:
            b.restart_probability = float(rng.choice([0.0, 1.0, rng.random()]))
    else:
        b.max_mne_level = int(rng.integers(b.mne_level, 4))
        if b.max_mne_level != b.mne_level:
            b.stagnation_step = int(rng.integers(1, 6))
        if rng.random() < 0.7:
            b.restart_stagnation, b.restart_cooldown = int(rng.integers(1, 12)), int(rng.integers(1, 6))
    return b


POLICY_FIELDS = ("kind", "mne_level", "max_mne_level", "region", "restart_mode", "restart_period",
                 "restart_probability", "stagnation_step", "restart_stagnation", "restart_cooldown")


def control_cases():
    rng = np.random.default_rng(4242)
    out = {"identity": [], "samples": [], "regions": [], "archive": [], "snapshot": [], "policy": [],
           "feedback": []}
    for n in (3, 5, 17, 64):
        for _ in range(6):
            a, b = rng.permutation(n).astype(np.uint32), rng.permutation(n).astype(np.uint32)
            out["identity"].append({"a": a.tolist(), "b": b.tolist(), "id": [str(v) for v in O.ref_fingerprint(a)],
                                    "compare": O.ref_canonical_compare(a, b),
                                    "compare_rotated": O.ref_canonical_compare(a, np.roll(a, 3)[::-1].copy())})
    for n in (3, 7, 16, 17, 64, 100, 1000):
        for _ in range(3):
            key = int(rng.integers(0, 2**63))
            out["samples"].append({"n": n, "key": str(key), "nodes": O.ref_sample_nodes(n, key).tolist()})
    for n in (3, 7, 17, 40, 200):
        for _ in range(3):
            tour = rng.permutation(n).astype(np.uint32)
            width = min(4, n - 1)
            primary = np.array([rng.permutation([j for j in range(n) if j != a])[:width] for a in range(n)],
                               dtype=np.uint32)
            if n > 10:
                primary[::5, -1] = n  # empty slots on the right of sparse rows
            key, batch = int(rng.integers(0, 2**63)), int(rng.integers(0, 1000))
            regions = [[O.ref_build_region(tour, primary, key, batch, m, r).tolist() for r in range(4)]
                       for m in range(2)]
            out["regions"].append({"tour": tour.tolist(), "primary": primary.tolist(), "key": str(key),
                                   "batch": batch, "regions": regions})
    for n in (5, 8, 12, 30):
        for rep in range(8):
            xy = rng.integers(0, 200 if rep % 2 else 10001, size=(n, 2))
            M = O.distance_matrix(xy)
            prep = O.prepare(xy, O.default_settings(primary_width=min(4, n - 1), backup_width=0,
                                                    ls_width=min(6, n - 1)))
            g = prep["initial_tour"].astype(np.uint32)
            pool = [g] + tour_variants(rng, g, 10) + [np.roll(g, 3).astype(np.uint32), g[::-1].copy()]
            costs = [int(M[t, np.roll(t, -1)].sum()) for t in pool]
            gi = int(np.argmin(costs))
            rest = [t for k, t in enumerate(pool) if k != gi]
            k = int(rng.integers(0, 4))
            chosen = O.ref_archive_update(M, pool[gi], np.array(rest[:k]).reshape(-1, n),
                                          np.array(rest[k:]).reshape(-1, n))
            out["archive"].append({"xy": xy.tolist(), "global": pool[gi].tolist(),
                                   "entries": [t.tolist() for t in rest[:k]],
                                   "completed": [t.tolist() for t in rest[k:]], "chosen": chosen.tolist()})
    for n in (3, 5, 9, 17, 40):
        for rep in range(6):
            xy = rng.integers(0, [10001, 30, 3][rep % 3], size=(n, 2))
            M = O.distance_matrix(xy)
            pw = min(int(rng.integers(2, 6)), n - 1)
            prep = O.prepare(xy, O.default_settings(primary_width=pw, backup_width=0, ls_width=min(6, n - 1)))
            g = prep["initial_tour"].astype(np.uint32)
            pool = [g] + tour_variants(rng, g, 6)
            costs = [int(M[t, np.roll(t, -1)].sum()) for t in pool]
            archive_in = np.array([pool[i] for i in np.argsort(costs, kind="stable")])
            active = pool[int(rng.integers(len(pool)))]
            key, batch = int(rng.integers(0, 2**63)), int(rng.integers(0, 1000))
            samples = O.ref_sample_nodes(n, key ^ 977)
            fb = O.Feedback(float(rng.random()), float(rng.random()), int(rng.integers(0, 40)), 3, 1)
            deps = np.array([rng.permutation(n) for _ in range(int(rng.integers(0, 3)))], dtype=np.uint32).reshape(-1, n)
            amounts = rng.random(deps.shape[0]) * 0.01
            lo, hi = sorted(float(v) for v in rng.random(2) * 0.01)
            if rep % 5 == 0:
                lo = hi
            mask = int(rng.choice([0xFFFFFFFF, 0xFFFF, 0x00FF00FF]))
            elapsed = float(rng.random() * 1.2)
            r = O.ref_action_snapshot(M, prep["primary"], prep["local_scale"], 1e-12, samples, active, archive_in,
                                      0.005, 0.7, 0.0001, deps, amounts, lo, hi, fb, elapsed, key, batch, mask)
            out["snapshot"].append({
                "xy": xy.tolist(), "primary": prep["primary"].tolist(), "local_scale": prep["local_scale"].tolist(),
                "samples": samples.tolist(), "active": active.tolist(), "archive": r["archive"].tolist(),
                "trails": [v.hex() for v in r["trails"].reshape(-1)], "default": r["default"].hex(),
                "minimum": lo.hex(), "maximum": hi.hex(),
                "feedback": [fb.return_rate, fb.ls_work, fb.stagnant_batches, fb.epoch_batches, fb.restarts],
                "elapsed": elapsed.hex(), "key": str(key), "batch": batch, "mask": mask,
                "features_bits": r["features"].view(np.uint32).reshape(-1).tolist(),
                "regions": r["regions"].tolist(), "alternative": r["alternative"], "legal_mask": r["mask"]})
    for _ in range(400):
        b = random_policy(rng)
        fb = O.Feedback(0, 0, int(rng.choice([0, 1, 5, 9, 30, 2**64 - 1])), int(rng.integers(0, 20)), 0)
        batch, u = int(rng.integers(0, 50)), float(rng.random())
        mask = int(rng.choice([0xFFFFFFFF, 0xFFFF, int(rng.integers(1, 2**32))])) | (1 << (b.region * 4 + b.mne_level))
        legal = int(rng.integers(1, 2**32)) | 1
        out["policy"].append({"policy": [getattr(b, f) for f in POLICY_FIELDS],
                              "feedback": [fb.stagnant_batches, fb.epoch_batches], "batch": batch, "u": u.hex(),
                              "mask": mask, "action": O.ref_baseline_action(b, fb, batch, u, mask), "legal": legal,
                              "factorial": [O.ref_factorial_mask(v, b, fb, batch, u, legal) for v in range(4)]})
    for _ in range(200):
        fb = O.Feedback(float(rng.random()), float(rng.random()), int(rng.integers(0, 5)), int(rng.integers(0, 5)), 0)
        before = [fb.return_rate, fb.ls_work, fb.stagnant_batches, fb.epoch_batches, fb.restarts]
        args = [bool(rng.integers(0, 2)), float(rng.random() * 1.5 - 0.2), float(rng.random() * 1.5 - 0.2),
                int(rng.integers(0, 3))]
        O.ref_update_feedback(fb, *args)
        restart = bool(rng.random() < 0.3)
        if restart:
            O.ref_restart_feedback(fb)
        out["feedback"].append({"before": before, "args": [args[0], args[1].hex(), args[2].hex(), args[3]],
                                "restart": restart,
                                "after_bits": [int(np.float32(fb.return_rate).view(np.uint32)),
                                               int(np.float32(fb.ls_work).view(np.uint32)),
                                               fb.stagnant_batches, fb.epoch_batches, fb.restarts]})
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` in a checkout with the reference")
    payload = {
        "construct_ls.json": construct_ls_cases(),
        "select_next.json": select_next_cases(),
        "trail_limits.json": trail_limit_cases(),
        "pheromone.json": pheromone_cases(),
        "two_opt.json": two_opt_cases(),
        "score_cpu.json": score_cases(),
        "control.json": control_cases(),
    }
    for name, data in payload.items():
        (OUT / name).write_text(json.dumps(data, separators=(",", ":")))
        print(name, (OUT / name).stat().st_size, "bytes")


if __name__ == "__main__":
    main()
