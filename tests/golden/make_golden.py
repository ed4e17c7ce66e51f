"""Write golden vectors from the reference's own CPU core (oracle/_ref).

Run in a checkout that has the reference sources (the recipe in
oracle/Makefile builds oracle/_ref/libgpfaco_ref.so from them):

    make -C oracle ref && python tests/golden/make_golden.py

Inputs are generated here from fixed seeds; expected outputs come from the
reference's FocusedConstruction + checklist_two_opt, select_next,
candidate_trail_limits, SparsePheromone and score_cpu, called with the
integer EUC_2D distances this engine uses. The fixtures are data only.
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
from tests.program_gen import random_program  # noqa: E402

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
    }
    for name, data in payload.items():
        (OUT / name).write_text(json.dumps(data, separators=(",", ":")))
        print(name, (OUT / name).stat().st_size, "bytes")


if __name__ == "__main__":
    main()
