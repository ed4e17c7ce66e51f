"""Synthetic Euclidean TSP instances with integer coordinates.

Uniform: integer coordinates uniform in [0, side]^2. Clustered: 8-32 Gaussian
centres uniform in the square, sigma = 2% of the side, points rounded and
clipped to the square. Every draw comes from a SplitMix64 stream keyed by
(seed, instance) so the same instance is produced on any machine without
depending on numpy's generator versions.
"""
from __future__ import annotations

import math

import numpy as np

_MASK = (1 << 64) - 1


def _splitmix(state: int):
    while True:
        state = (state + 0x9E3779B97F4A7C15) & _MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        yield z ^ (z >> 31)


def _uniform_words(seed: int, count: int) -> np.ndarray:
    """`count` 64-bit SplitMix64 words (vectorised)."""
    i = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def _key(seed: int, instance: int, salt: int) -> int:
    gen = _splitmix((seed * 0x100000001B3 + instance * 0x9E3779B1 + salt) & _MASK)
    return next(gen)


def uniform_instance(n: int, seed: int = 0, instance: int = 0, side: int = 10000) -> np.ndarray:
    """Integer coordinates uniform in [0, side]^2 as float64 [n, 2]."""
    words = _uniform_words(_key(seed, instance, 1), 2 * n)
    values = (words % np.uint64(side + 1)).astype(np.float64)
    return values.reshape(n, 2)


def clustered_instance(n: int, seed: int = 0, instance: int = 0, side: int = 10000) -> np.ndarray:
    """8-32 Gaussian clusters, sigma = 2% of the side, rounded and clipped."""
    words = _uniform_words(_key(seed, instance, 2), 4 * n + 128)
    u = (words >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    clusters = 8 + int(u[0] * 25)
    centres = u[1: 1 + 2 * clusters].reshape(clusters, 2) * side
    rest = u[1 + 2 * clusters:]
    member = np.minimum((rest[:n] * clusters).astype(np.int64), clusters - 1)
    u1 = np.maximum(rest[n: 2 * n], 2.0 ** -53)
    u2 = rest[2 * n: 3 * n]
    radius = np.sqrt(-2.0 * np.log(u1)) * (0.02 * side)
    points = centres[member] + np.stack([radius * np.cos(2 * math.pi * u2), radius * np.sin(2 * math.pi * u2)], 1)
    return np.clip(np.rint(points), 0, side).astype(np.float64)


def instance_key(n: int, seed: int, instance: int, clustered: bool = False) -> int:
    """Stable 64-bit key of a generated instance."""
    return _key(seed, instance, 3 + (n << 8) + (1 if clustered else 0)) & ((1 << 63) - 1)
