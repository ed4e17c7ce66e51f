"""Learned edge prior: training data from the engine's own best tours, the
MLP on the HIP matrix cores, and its use as FACO heuristic multipliers.

The reference has no learned prior. The network (8 -> 64 -> 64 -> 1, tanh
GELU) scores each primary candidate edge from eight features; sigmoid of
the score multiplies the edge's heuristic, so FACO itself is unchanged and
plain FACO is the same engine with no prior set. Training labels are the
edges of the best tours plain FACO finds on synthetic instances.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import time
from pathlib import Path

import numpy as np

from ._native import check, lib

PARAM_COUNT = 4801
# Weights trained by `python -m gplsaco_amd.prior train` (k = 20 candidates).
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "data" / "edge_prior_k20.npz"
# Multipliers are sigmoid(logit) ** PRIOR_EXPONENT. 2 was chosen among
# 1, 1.5, 2, 2.5, 3 and 4 on validation instances (instance seed 4242, 16
# TSP1000 instances x 8 seeds, 300 iterations), not on the comparison
# instances (seed 7919); with labels from 1000-iteration 2-opt + Or-opt runs
# it stays the best of 1.5, 2 and 3 on the same validation instances.
PRIOR_EXPONENT = 2.0


class EdgePrior:
    """The network's weights on the device (FP32 master + Adam state)."""

    def __init__(self, seed: int = 0, params=None):
        handle = ctypes.c_void_p()
        check(lib().faco_prior_create(int(seed), ctypes.byref(handle)))
        self._handle = handle
        if params is not None:
            self.set_params(params)

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            lib().faco_prior_destroy(handle)
            self._handle = None

    def params(self) -> np.ndarray:
        out = np.zeros(PARAM_COUNT, dtype=np.float32)
        check(lib().faco_prior_get_params(self._handle, out))
        return out

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != PARAM_COUNT:
            raise ValueError(f"the prior has {PARAM_COUNT} parameters")
        check(lib().faco_prior_set_params(self._handle, p))

    def save(self, path, info=None) -> None:
        """Parameters, layer widths and a JSON description of the training run."""
        np.savez(path, params=self.params(), spec=np.array([8, 64, 64, 1]),
                 info=np.array(json.dumps(info or {})))

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS) -> "EdgePrior":
        with np.load(path, allow_pickle=False) as data:
            if data["spec"].tolist() != [8, 64, 64, 1]:
                raise ValueError("not an 8-64-64-1 edge prior")
            prior = cls(params=data["params"])
            prior.info = json.loads(str(data["info"])) if "info" in data else {}
            return prior

    def logits(self, features) -> np.ndarray:
        f = _features(features)
        out = np.zeros(f.shape[0], dtype=np.float32)
        check(lib().faco_prior_logits(self._handle, f.reshape(-1), f.shape[0], out))
        return out

    def gradients(self, features, labels, pos_weight: float = 1.0):
        f, y = _features(features), _labels(labels, features)
        g = np.zeros(PARAM_COUNT, dtype=np.float32)
        loss = ctypes.c_float()
        check(lib().faco_prior_gradients(self._handle, f.reshape(-1), y, f.shape[0], float(pos_weight), g,
                                         ctypes.byref(loss)))
        return loss.value, g

    def train(self, features, labels, *, batch: int = 8192, epochs: int = 4, lr: float = 3e-3,
              pos_weight: float = 4.0, seed: int = 0) -> np.ndarray:
        """Adam epochs over shuffled minibatches; mean batch loss per epoch."""
        f, y = _features(features), _labels(labels, features)
        losses = np.zeros(epochs, dtype=np.float32)
        check(lib().faco_prior_train(self._handle, f.reshape(-1), y, f.shape[0], batch, epochs, float(lr),
                                     float(pos_weight), int(seed), losses))
        return losses


def mma_check(A, B) -> np.ndarray:
    """One 16x16x32 bf16 matrix-core product A[16,32] @ B[32,16] on the device."""
    a = np.ascontiguousarray(A, dtype=np.float32)
    b = np.ascontiguousarray(B, dtype=np.float32)
    c = np.zeros((16, 16), dtype=np.float32)
    check(lib().faco_prior_mma_check(a.reshape(-1), b.reshape(-1), c.reshape(-1)))
    return c


def _features(features) -> np.ndarray:
    f = np.ascontiguousarray(features, dtype=np.float32)
    if f.ndim != 2 or f.shape[1] != 8 or f.shape[0] == 0:
        raise ValueError("features must be [edges, 8]")
    return f


def _labels(labels, features) -> np.ndarray:
    y = np.ascontiguousarray(labels, dtype=np.uint8).reshape(-1)
    if y.size != np.asarray(features).shape[0]:
        raise ValueError("one label per edge")
    return y


def tour_labels(primary, tour) -> np.ndarray:
    """1 for the candidate edges (a, primary[a][j]) that are edges of `tour`."""
    primary = np.asarray(primary)
    n = primary.shape[0]
    tour = np.asarray(tour, dtype=np.int64)
    succ = np.empty(n, dtype=np.int64)
    pred = np.empty(n, dtype=np.int64)
    succ[tour] = np.roll(tour, -1)
    pred[tour] = np.roll(tour, 1)
    return ((primary == succ[:, None]) | (primary == pred[:, None])).astype(np.uint8).reshape(-1)


def training_instances(count: int, seed: int = 0, n_range=(200, 1000), size_step: int = 100):
    """(n, clustered, index) of `count` training instances: n uniform in
    `n_range` (SplitMix64 words of `seed`) rounded to a multiple of
    `size_step`, alternately uniform and clustered."""
    from .instances import _key, _uniform_words
    lo, hi = n_range
    words = _uniform_words(_key(seed, 0, 0x51A5), count)
    u = (words >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    sizes = np.clip(np.rint((lo + u * (hi - lo)) / size_step) * size_step, lo, hi).astype(int)
    return [(int(sizes[i]), i % 2 == 1, i) for i in range(count)]


def synthetic_instance(n, clustered, index, seed):
    """(instance key, coordinates) of a generated instance."""
    from .instances import clustered_instance, instance_key, uniform_instance
    gen = clustered_instance if clustered else uniform_instance
    return instance_key(n, seed, index, clustered), gen(n, seed=seed, instance=index)


def build_dataset(count: int, *, iterations: int = 200, primary_width: int = 20, seed: int = 0,
                  n_range=(200, 1000), size_step: int = 100, local_search: str = "2opt", log=None):
    """Training edges of `count` synthetic instances (training_instances):
    each instance is solved by plain FACO (one colony, `iterations`
    iterations, default ants, `local_search` "2opt" or "2opt+oropt") on the
    GPU, and its candidate edges are labelled by membership in the colony's
    best tour. Returns (features [E, 8] float32, labels [E] uint8, per-size
    summary)."""
    from . import gp_faco_ext as ext
    groups = {}
    for n, clustered, index in training_instances(count, seed, n_range, size_step):
        groups.setdefault(n, []).append((clustered, index))
    feats, labs, summary = [], [], []
    for n in sorted(groups):
        items = groups[n]
        settings = ext.FixedFacoSettings()
        settings.primary_width = primary_width
        engine = ext.FacoBatchEngine(n, len(items), settings)
        if local_search != "2opt":
            engine.set_local_search(local_search)
        keys = []
        for clustered, index in items:
            key, xy = synthetic_instance(n, clustered, index, seed)
            engine.register_problem(key, xy)
            keys.append(key)
        ants = engine.settings.ants
        out = engine.evaluate_faco_evaluations(np.array(keys, dtype=np.uint64), np.full(len(keys), seed, np.uint64),
                                               ants * iterations)
        for key, item in zip(keys, out["items"]):
            primary = engine.problem_data(key)["primary"]
            feats.append(engine.edge_features(key))
            labs.append(tour_labels(primary, item["tour"]))
        summary.append({"n": n, "instances": len(items), "search_seconds": out["search_seconds"],
                        "mean_cost": float(np.mean([i["cost"] for i in out["items"]]))})
        if log:
            log(f"dataset n={n}: {len(items)} instances, {out['search_seconds']:.2f} s")
        del engine
    return np.concatenate(feats), np.concatenate(labs), summary


def multipliers(logits, exponent: float = PRIOR_EXPONENT) -> np.ndarray:
    """Heuristic multipliers sigmoid(logit) ** exponent (float32, > 0)."""
    s = 1.0 / (1.0 + np.exp(-np.asarray(logits, dtype=np.float64)))
    return np.maximum(s ** exponent, 1e-30).astype(np.float32)


def equal_budget_comparison(prior: "EdgePrior", *, n: int = 1000, instances: int = 8, seeds: int = 4,
                            iterations: int = 100, fractions=None, primary_width: int = 20,
                            exponent: float = PRIOR_EXPONENT,
                            seed: int = 7919):
    """Learned-prior FACO against plain FACO on unseen instances (half
    uniform, half clustered; instance seed `seed`, not a training seed).

    Plain FACO runs `iterations` iterations per colony; the learned-prior
    runs use each fraction of that budget (the same seeds, so every shorter
    run is a prefix of the longer one). Returns plain's mean final tour
    length, the learned mean at each fraction, the gap at the full budget
    ((learned - plain) / plain) and the smallest fraction whose learned mean
    reaches plain's mean (None when none does)."""
    from . import gp_faco_ext as ext
    fractions = sorted(fractions or [round(0.05 * k, 2) for k in range(4, 21)])
    settings = ext.FixedFacoSettings()
    settings.primary_width = primary_width
    colonies = instances * seeds
    engine = ext.FacoBatchEngine(n, colonies, settings)
    keys = []
    for i in range(instances):
        key, xy = synthetic_instance(n, i % 2 == 1, i, seed)
        engine.register_problem(key, xy)
        keys.append(key)
    task_keys = np.array([keys[c % instances] for c in range(colonies)], dtype=np.uint64)
    task_seeds = np.array([1 + c // instances for c in range(colonies)], dtype=np.uint64)
    ants = engine.settings.ants
    mean = lambda out: float(np.mean([it["cost"] for it in out["items"]]))  # noqa: E731
    plain = mean(engine.evaluate_faco_evaluations(task_keys, task_seeds, ants * iterations))
    for key in keys:
        engine.set_edge_prior(key, multipliers(prior.logits(engine.edge_features(key)), exponent))
    learned = {}
    for f in fractions:
        its = max(1, int(round(f * iterations)))
        learned[f] = mean(engine.evaluate_faco_evaluations(task_keys, task_seeds, ants * its))
    reach = next((f for f in fractions if learned[f] <= plain), None)
    full = learned[fractions[-1]]
    return {"n": n, "instances": instances, "seeds": seeds, "iterations": iterations, "ants": ants,
            "primary_width": primary_width, "exponent": exponent, "plain_mean": plain,
            "learned_mean": {str(f): v for f, v in learned.items()}, "gap_at_equal_fe": (full - plain) / plain,
            "fe_fraction_to_reach_plain": reach}


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(prog="python -m gplsaco_amd.prior",
                                     description="Train the learned edge prior on the GPU, or compare it with plain FACO.")
    sub = parser.add_subparsers(dest="command", required=True)
    t = sub.add_parser("train")
    t.add_argument("--instances", type=int, default=2000)
    t.add_argument("--iterations", type=int, default=1000, help="FACO iterations per labelling run")
    t.add_argument("--local-search", default="2opt+oropt", choices=["2opt", "2opt+oropt"],
                   help="local search of the labelling runs")
    t.add_argument("--epochs", type=int, default=8)
    t.add_argument("--batch", type=int, default=8192)
    t.add_argument("--lr", type=float, default=5e-4)
    t.add_argument("--pos-weight", type=float, default=1.0)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    c = sub.add_parser("compare")
    c.add_argument("--weights", default=str(DEFAULT_WEIGHTS))
    c.add_argument("--iterations", type=int, default=300)
    c.add_argument("--instances", type=int, default=16)
    c.add_argument("--seeds", type=int, default=8)
    c.add_argument("--exponent", type=float, default=PRIOR_EXPONENT)
    args = parser.parse_args(argv)
    if args.command == "train":
        started = time.perf_counter()
        f, y, summary = build_dataset(args.instances, iterations=args.iterations, seed=args.seed,
                                      local_search=args.local_search, log=lambda m: print(m, flush=True))
        built = time.perf_counter()
        prior = EdgePrior(seed=args.seed + 1)
        losses = prior.train(f, y, batch=args.batch, epochs=args.epochs, lr=args.lr, pos_weight=args.pos_weight,
                             seed=args.seed + 3)
        info = {"instances": args.instances, "labelling_iterations": args.iterations,
                "labelling_local_search": args.local_search, "edges": int(f.shape[0]),
                "positive_rate": float(y.mean()), "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
                "pos_weight": args.pos_weight, "seed": args.seed, "epoch_losses": [float(v) for v in losses],
                "dataset_seconds": built - started, "training_seconds": time.perf_counter() - built,
                "sizes": summary}
        Path(args.out).parent.mkdir(parents=True, exist_ok=True)
        prior.save(args.out, info)
        print(json.dumps({k: v for k, v in info.items() if k != "sizes"}), flush=True)
    else:
        prior = EdgePrior.load(args.weights)
        print(json.dumps(equal_budget_comparison(prior, iterations=args.iterations, instances=args.instances,
                                                 seeds=args.seeds, exponent=args.exponent)), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
