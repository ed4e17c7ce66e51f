"""The oracle of the learned edge prior (the reference has none), pinned by
what the mathematics fixes:

- bf16 rounding against torch's own float32 -> bfloat16 conversion;
- the FP64 network's gradient against central finite differences, for every
  parameter block (a dropped term, a wrong sign or a transposed operand in
  the backward pass fails it);
- the loss and output-bias gradient at zero weights in closed form
  (softplus(0) = log 2, sigmoid(0) = 1/2), which pins the positive weight;
- the rounded network (the device's rounding points) close to the FP64 one;
- the edge features by properties that hold for every instance: translation
  invariance, the column index, the mutual flag and rank read back from the
  neighbour's row, the shared-entry count, and the distance ratios;
- tour labels against a loop over the tour's edges.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import prior_oracle as PO
from gplsaco_amd.instances import clustered_instance, uniform_instance


def _params(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(PO.P) * scale).astype(np.float32)


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, -0.0, 65504.0,
                                  1.0 + 2.0 ** -9], dtype=np.float32)]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(PO.bf16(x), want)


BLOCKS = {"W1": (PO.O_W1, PO.O_B1), "b1": (PO.O_B1, PO.O_W2), "W2": (PO.O_W2, PO.O_B2), "b2": (PO.O_B2, PO.O_W3),
          "w3": (PO.O_W3, PO.O_B3), "b3": (PO.O_B3, PO.P)}


@pytest.mark.parametrize("pos_weight", [1.0, 3.5])
def test_fp64_gradient_matches_finite_differences(pos_weight):
    rng = np.random.default_rng(7)
    p = _params(3).astype(np.float64)
    x = rng.uniform(0, 4, size=(37, 8))
    y = (rng.random(37) < 0.3).astype(np.uint8)
    _, g = PO.loss_and_grad(p, x, y, pos_weight, rounding=False)
    h = 1e-6
    for name, (lo, hi) in BLOCKS.items():
        idx = rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False)
        for i in idx:
            q = p.copy(); q[i] += h
            lp, _ = PO.loss_and_grad(q, x, y, pos_weight, rounding=False)
            q[i] -= 2 * h
            lm, _ = PO.loss_and_grad(q, x, y, pos_weight, rounding=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (name, i, fd, g[i])


def test_loss_and_bias_gradient_at_zero_weights():
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 4, size=(50, 8))
    y = (rng.random(50) < 0.4).astype(np.uint8)
    w = np.where(y > 0, 3.0, 1.0)
    for rounding in (False, True):
        loss, g = PO.loss_and_grad(np.zeros(PO.P), x, y, 3.0, rounding=rounding)
        assert loss == pytest.approx(float(np.mean(w)) * np.log(2.0), rel=1e-12)
        assert g[PO.O_B3] == pytest.approx(float(np.mean(w * (0.5 - y))), rel=1e-6)
        # with zero weights every hidden unit is gelu(0) = 0: no other gradient
        assert np.count_nonzero(g[:PO.O_B3]) == 0


def test_rounded_network_is_close_to_fp64():
    rng = np.random.default_rng(4)
    p = _params(5, 0.25)
    x = rng.uniform(0, 4, size=(400, 8)).astype(np.float32)
    y = (rng.random(400) < 0.25).astype(np.uint8)
    s_r = PO.forward(p, x, rounding=True)["s"]
    s_e = PO.forward(p, x, rounding=False)["s"]
    assert np.max(np.abs(s_r - s_e)) < 0.05 * (1 + np.max(np.abs(s_e)))
    _, g_r = PO.loss_and_grad(p, x, y, 2.0, rounding=True)
    _, g_e = PO.loss_and_grad(p, x, y, 2.0, rounding=False)
    for name, (lo, hi) in BLOCKS.items():
        a, b = g_r[lo:hi], g_e[lo:hi]
        assert np.linalg.norm(a - b) <= 0.03 * np.linalg.norm(b) + 1e-9, name


def _prepared(xy):
    s = O.default_settings()
    P = O.prepare(xy, s)
    span = np.ptp(np.asarray(xy, dtype=np.float64), axis=0)
    eps = 1e-12 * max(span[0], span[1], 1.0)
    return P, eps


@pytest.mark.parametrize("kind,n", [("uniform", 150), ("clustered", 400), ("duplicates", 60)])
def test_edge_feature_properties(kind, n):
    if kind == "clustered":
        xy = clustered_instance(n, seed=2)
    else:
        xy = uniform_instance(n, seed=2)
        if kind == "duplicates":
            xy[5::4] = xy[1]
    P, eps = _prepared(xy)
    primary, scale = P["primary"], P["local_scale"]
    pw = primary.shape[1]
    f = O.prior_features(xy, primary, scale, eps).reshape(n, pw, 8)
    M = O.distance_matrix(xy)
    assert np.all(np.isfinite(f)) and f.min() >= 0
    np.testing.assert_array_equal(f[:, :, 1], np.broadcast_to(np.arange(pw, dtype=np.float32) / np.float32(pw - 1),
                                                              (n, pw)))
    sets = [set(r.tolist()) for r in primary]
    mean_scale = max(float(np.sum(scale)) / n, eps)
    for a in range(n):
        for j in range(pw):
            b = int(primary[a, j])
            brow = primary[b].tolist()
            mutual = a in brow
            assert f[a, j, 3] == (1.0 if mutual else 0.0)
            rank = brow.index(a) if mutual else pw
            assert f[a, j, 4] == np.float32(rank) / np.float32(pw)
            assert round(float(f[a, j, 7]) * pw) == len(sets[a] & set(brow))
            d = float(M[a, b])
            for col, denom in ((0, max(scale[a], eps)), (5, max(scale[b], eps)), (6, mean_scale),
                               (2, max(float(M[a, primary[a, 0]]), 1.0))):
                assert f[a, j, col] == pytest.approx(min(d / denom, 4.0), rel=1e-6, abs=1e-30), (a, j, col)
    # rows are sorted by distance, so the nearest-distance ratio never falls
    assert np.all(np.diff(f[:, :, 2], axis=1) >= 0)
    moved = np.asarray(xy) + np.array([123, 77])
    P2, eps2 = _prepared(moved)
    np.testing.assert_array_equal(P2["primary"], primary)
    np.testing.assert_array_equal(O.prior_features(moved, P2["primary"], P2["local_scale"], eps2).reshape(n, pw, 8), f)


def test_tour_labels_against_edge_loop():
    from gplsaco_amd.prior import tour_labels
    rng = np.random.default_rng(9)
    n, pw = 40, 6
    primary = np.array([rng.choice([b for b in range(n) if b != a], pw, replace=False) for a in range(n)])
    tour = rng.permutation(n)
    edges = {frozenset((int(tour[i]), int(tour[(i + 1) % n]))) for i in range(n)}
    want = np.array([frozenset((a, int(primary[a, j]))) in edges for a in range(n) for j in range(pw)], dtype=np.uint8)
    np.testing.assert_array_equal(tour_labels(primary, tour), want)
    np.testing.assert_array_equal(PO.labels_from_tour(primary, tour), want)


def test_prior_inputs_are_validated_on_the_host():
    from gplsaco_amd import prior
    with pytest.raises(ValueError):
        prior._features(np.zeros((4, 7)))
    with pytest.raises(ValueError):
        prior._features(np.zeros((0, 8)))
    with pytest.raises(ValueError):
        prior._labels(np.zeros(3), np.zeros((4, 8)))
