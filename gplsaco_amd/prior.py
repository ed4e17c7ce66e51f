"""Learned edge prior: training data from the engine's own best tours, the
MLP on the HIP matrix cores, and its use as FACO heuristic multipliers.

The reference has no learned prior. The network (8 -> 64 -> 64 -> 1, tanh
GELU) scores each primary candidate edge from eight features; sigmoid of
the score multiplies the edge's heuristic, so FACO itself is unchanged and
plain FACO is the same engine with no prior set. Training labels are the
edges of the best tours plain FACO finds on synthetic instances.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._native import check, lib

PARAM_COUNT = 4801


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

    def save(self, path) -> None:
        np.savez(path, params=self.params(), spec=np.array([8, 64, 64, 1]))

    @classmethod
    def load(cls, path) -> "EdgePrior":
        with np.load(path, allow_pickle=False) as data:
            return cls(params=data["params"])

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


def build_dataset(sizes, per_size: int, iterations: int, seed: int = 0):
    """Features and labels from plain FACO's best tours on synthetic instances:
    `per_size` instances (alternately uniform and clustered) of each size,
    `iterations` FACO iterations each. Returns (features [E, 8], labels [E],
    per-size summary)."""
    from . import gp_faco_ext as ext
    from .instances import clustered_instance, instance_key, uniform_instance
    feats, labs, summary = [], [], []
    for n in sizes:
        engine = ext.FacoBatchEngine(n, per_size, ext.FixedFacoSettings())
        keys = []
        for i in range(per_size):
            clustered = i % 2 == 1
            gen = clustered_instance if clustered else uniform_instance
            key = instance_key(n, seed, i, clustered)
            engine.register_problem(key, gen(n, seed=seed, instance=i))
            keys.append(key)
        ants = engine.settings.ants
        out = engine.evaluate_faco_evaluations(np.array(keys, dtype=np.uint64), np.full(per_size, seed, np.uint64),
                                               ants * iterations)
        for key, item in zip(keys, out["items"]):
            primary = engine.problem_data(key)["primary"]
            feats.append(engine.edge_features(key))
            labs.append(tour_labels(primary, item["tour"]))
        summary.append({"n": n, "instances": per_size, "mean_cost": float(np.mean([i["cost"] for i in out["items"]]))})
    return np.concatenate(feats), np.concatenate(labs), summary
