"""Known-answer checks of the oracle's batch selection (reduce_and_select,
reference cuda/colony_state.cuh:148-208) and restart transaction
(apply_control_action, reference cuda/control_state.cuh:309-364) on explicit,
hand-made colony states.

Each case is chosen so that a plausible slip fails it: the epoch-source coin
read the wrong way round, iteration-best ties going to a later ant, the MMAS
bounds taken from the global or the iteration cost instead of the epoch cost,
the epoch or global record replaced without a strict improvement, the restart
leaving the global record changed, taking the bounds from anything but the
archive entry, or skipping the feedback reset. The pheromone after a batch is
compared with the reference's own SparsePheromone (oracle/_ref) where that is
built, and with the evaporate-then-deposit rule of the reference's replay test
(tests/cpp/test_fixed_faco.cpp:110-118) otherwise.
"""
import numpy as np
import pytest

from oracle import oracle as O

N = 14


def _instance(seed=3):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 10001, size=(N, 2)).astype(np.int32)


def _settings(probability):
    return O.default_settings(ants=4, primary_width=5, backup_width=3, ls_width=6, retention=0.75,
                              epoch_source_probability=probability)


def _tours(seed, count):
    rng = np.random.default_rng(seed)
    return np.stack([rng.permutation(N) for _ in range(count)]).astype(np.uint32)


def _state(prep, epoch_cost, global_cost, parent_cost, seed=11):
    """A colony whose parent, epoch and global tours are three distinct
    permutations, with uniform trails at the initial maximum."""
    parent, epoch, glob = _tours(seed, 3)
    pw = prep["settings"].primary_width
    trail = O.trail_limits(pw, 0.1, 0.75, float(prep["initial_cost"]))[1]
    return dict(costs=(epoch_cost, global_cost, parent_cost), tours=np.stack([parent, epoch, glob]),
                trails=np.full((N, pw), trail), default_trail=trail)


def _expected_trails(xy, prep, parent, trail, minimum, maximum, parent_cost, retention):
    primary = prep["primary"]
    if O.ref_available():
        M = O.distance_matrix(xy).astype(np.float64)
        trails, defaults, _ = O.ref_pheromone_run(M, primary, trail, parent[None], retention, [minimum], [maximum],
                                                  [1.0 / parent_cost])
        return trails[0].reshape(primary.shape), float(defaults[0])
    pos = np.argsort(parent)
    out = np.empty(primary.shape)
    for a in range(N):
        neighbours = (parent[(pos[a] + 1) % N], parent[(pos[a] - 1) % N])
        for j in range(primary.shape[1]):
            v = max(minimum, trail * retention)
            if primary[a, j] in neighbours:
                v = min(maximum, v + 1.0 / parent_cost)
            out[a, j] = v
    return out, max(minimum, trail * retention)


def _select(xy, s, state, ant_tours, ant_costs, key=77, batch=5):
    return O.colony_select(xy, s, key, batch, ant_tours, ant_costs, state["costs"], state["tours"], state["trails"],
                           state["default_trail"])


def _coin(probability, key=77, batch=5):
    # the epoch-source coin: ant 0xffffffff, step 0 of the colony's stream
    return O.uniform53(key, batch, 0xFFFFFFFF, 0) < probability


@pytest.mark.parametrize("ant_costs,best", [((50, 30, 30, 40), 1), ((9, 5, 8, 5), 1), ((7, 7, 7, 7), 0),
                                            ((3, 4, 5, 6), 0), ((6, 5, 4, 3), 3)])
def test_iteration_best_is_the_first_minimum(ant_costs, best):
    xy = _instance()
    s = _settings(0.0)
    prep = O.prepare(xy, s)
    state = _state(prep, 100, 100, 100)
    ants = _tours(21, 4)
    got = _select(xy, s, state, ants, ant_costs)
    assert got["iteration_best"] == best
    assert not got["source_is_epoch"]
    np.testing.assert_array_equal(got["parent"], ants[best])
    assert got["costs"][2] == ant_costs[best]


def test_probability_zero_takes_the_iteration_best_even_when_the_epoch_is_better():
    xy = _instance()
    s = _settings(0.0)
    prep = O.prepare(xy, s)
    pw = prep["settings"].primary_width
    state = _state(prep, epoch_cost=20, global_cost=15, parent_cost=20)
    ants = _tours(22, 4)
    got = _select(xy, s, state, ants, (40, 30, 35, 31))
    # no improvement: epoch and global tours and costs are kept
    np.testing.assert_array_equal(got["epoch"], state["tours"][1])
    np.testing.assert_array_equal(got["global"], state["tours"][2])
    assert tuple(got["costs"]) == (20, 15, 30)
    np.testing.assert_array_equal(got["parent"], ants[1])
    # bounds from the epoch cost: not the global (15) nor the iteration best (30)
    lo, hi = O.trail_limits(pw, s.p_best, s.retention, 20.0)
    assert (got["minimum"], got["maximum"]) == (lo, hi)
    trails, default = _expected_trails(xy, prep, ants[1], state["default_trail"], lo, hi, 30, s.retention)
    np.testing.assert_array_equal(got["trails"], trails)
    assert got["default_trail"] == default


def test_probability_one_takes_the_epoch_tour():
    xy = _instance()
    s = _settings(1.0)
    prep = O.prepare(xy, s)
    pw = prep["settings"].primary_width
    state = _state(prep, epoch_cost=20, global_cost=15, parent_cost=25)
    ants = _tours(23, 4)
    got = _select(xy, s, state, ants, (40, 30, 35, 31))
    assert got["source_is_epoch"] and got["source_uniform"] < 1.0
    np.testing.assert_array_equal(got["parent"], state["tours"][1])
    assert tuple(got["costs"]) == (20, 15, 20)
    lo, hi = O.trail_limits(pw, s.p_best, s.retention, 20.0)
    assert (got["minimum"], got["maximum"]) == (lo, hi)
    trails, _ = _expected_trails(xy, prep, state["tours"][1], state["default_trail"], lo, hi, 20, s.retention)
    np.testing.assert_array_equal(got["trails"], trails)


def test_strict_improvement_updates_epoch_and_global_separately():
    xy = _instance()
    s = _settings(1.0)
    prep = O.prepare(xy, s)
    pw = prep["settings"].primary_width
    ants = _tours(24, 4)
    # improves the epoch (40 -> 30) but not the global (25)
    state = _state(prep, epoch_cost=40, global_cost=25, parent_cost=40)
    got = _select(xy, s, state, ants, (33, 30, 30, 45))
    np.testing.assert_array_equal(got["epoch"], ants[1])
    np.testing.assert_array_equal(got["global"], state["tours"][2])
    np.testing.assert_array_equal(got["parent"], ants[1])
    assert tuple(got["costs"]) == (30, 25, 30)
    assert got["maximum"] == O.trail_limits(pw, s.p_best, s.retention, 30.0)[1]
    # an equal cost is not an improvement
    state = _state(prep, epoch_cost=30, global_cost=30, parent_cost=30)
    got = _select(xy, s, state, ants, (33, 30, 30, 45))
    np.testing.assert_array_equal(got["epoch"], state["tours"][1])
    np.testing.assert_array_equal(got["global"], state["tours"][2])
    np.testing.assert_array_equal(got["parent"], state["tours"][1])
    # improves both
    state = _state(prep, epoch_cost=90, global_cost=80, parent_cost=90)
    got = _select(xy, s, state, ants, (33, 31, 30, 45))
    np.testing.assert_array_equal(got["epoch"], ants[2])
    np.testing.assert_array_equal(got["global"], ants[2])
    assert tuple(got["costs"]) == (30, 30, 30)


@pytest.mark.parametrize("probability", [0.01, 0.3, 0.5, 0.99])
def test_coin_reads_the_colony_stream(probability):
    """Over many batches the coin decides exactly as u53(key, batch,
    0xffffffff, 0) < p, in both directions."""
    xy = _instance()
    s = _settings(probability)
    prep = O.prepare(xy, s)
    ants = _tours(25, 4)
    seen = set()
    for batch in range(40):
        state = _state(prep, epoch_cost=20, global_cost=20, parent_cost=20)
        got = _select(xy, s, state, ants, (40, 30, 35, 31), key=1234 + batch, batch=batch)
        want = _coin(probability, key=1234 + batch, batch=batch)
        assert got["source_is_epoch"] == want
        assert got["source_uniform"] == O.uniform53(1234 + batch, batch, 0xFFFFFFFF, 0)
        np.testing.assert_array_equal(got["parent"], state["tours"][1] if want else ants[1])
        assert got["costs"][2] == (20 if want else 30)
        seen.add(want)
    if 0.05 < probability < 0.95:
        assert seen == {True, False}


def _archive(prep, count, seed=31):
    tours = _tours(seed, count)
    costs = np.array([int(O.distance_matrix(prep["xy"])[t, np.roll(t, -1)].sum()) for t in tours], dtype=np.int64)
    return tours, costs


@pytest.mark.parametrize("alternative", [0, 1, 3])
@pytest.mark.parametrize("initializing", [False, True])
def test_restart_transaction(alternative, initializing):
    xy = _instance()
    s = _settings(0.01)
    prep = O.prepare(xy, s)
    prep["xy"] = xy
    pw = prep["settings"].primary_width
    archive, archive_costs = _archive(prep, 4)
    state = _state(prep, epoch_cost=12345, global_cost=11111, parent_cost=13000)
    # trails that differ per entry, so a restart that leaves any entry fails
    state["trails"] = np.linspace(1e-6, 2e-5, N * pw).reshape(N, pw)
    state["default_trail"] = 3e-6
    fb = O.Feedback(return_rate=0.4, ls_work=0.25, stagnant_batches=7, epoch_batches=9, restarts=2)
    got = O.control_restart(xy, s, archive, archive_costs, alternative, initializing, state["costs"],
                            state["tours"], state["trails"], state["default_trail"], fb)
    cost = int(archive_costs[alternative])
    # parent = epoch = the archive entry, with its cost; the global record stays
    np.testing.assert_array_equal(got["parent"], archive[alternative])
    np.testing.assert_array_equal(got["epoch"], archive[alternative])
    np.testing.assert_array_equal(got["global"], state["tours"][2])
    assert tuple(got["costs"]) == (cost, 11111, cost)
    # bounds from the entry's cost; every trail and the default at the maximum
    lo, hi = O.trail_limits(pw, s.p_best, s.retention, float(cost))
    assert (got["minimum"], got["maximum"]) == (lo, hi)
    assert np.all(got["trails"] == hi) and got["default_trail"] == hi
    M = O.distance_matrix(xy)
    eta = np.array([[1.0 / M[a, b] if M[a, b] else 1.0 for b in prep["primary"][a]] for a in range(N)])
    np.testing.assert_array_equal(got["products"], hi * eta)
    # the active and epoch identities become the entry's
    ident = tuple(int(v) for v in O.tour_ident(archive[alternative]))
    assert got["active_identity"] == ident and got["epoch_identity"] == ident
    # restart_feedback unless the transaction initialises the colony
    if initializing:
        assert (fb.return_rate, fb.ls_work, fb.epoch_batches, fb.restarts) == (np.float32(0.4), 0.25, 9, 2)
    else:
        assert (fb.return_rate, fb.ls_work, fb.epoch_batches, fb.restarts) == (0.0, 0.0, 0, 3)
    assert fb.stagnant_batches == 7


def test_restart_then_selection_takes_bounds_from_the_restored_epoch():
    """After a restart to a worse archive entry (global stays better), the next
    batch's bounds come from the restored epoch cost, and with p = 1 its
    parent is the restored tour."""
    xy = _instance()
    s = _settings(1.0)
    prep = O.prepare(xy, s)
    prep["xy"] = xy
    pw = prep["settings"].primary_width
    archive, archive_costs = _archive(prep, 2)
    glob = int(min(archive_costs)) - 100
    state = _state(prep, epoch_cost=glob, global_cost=glob, parent_cost=glob)
    fb = O.Feedback()
    got = O.control_restart(xy, s, archive, archive_costs, 1, False, state["costs"], state["tours"], state["trails"],
                            state["default_trail"], fb)
    restored = dict(costs=tuple(got["costs"]), tours=np.stack([got["parent"], got["epoch"], got["global"]]),
                    trails=got["trails"], default_trail=got["default_trail"])
    ants = _tours(26, 4)
    worse = int(archive_costs[1]) + 50
    nxt = _select(xy, s, restored, ants, (worse, worse + 1, worse + 2, worse + 3))
    np.testing.assert_array_equal(nxt["parent"], archive[1])
    assert nxt["maximum"] == O.trail_limits(pw, s.p_best, s.retention, float(archive_costs[1]))[1]
    assert tuple(nxt["costs"]) == (int(archive_costs[1]), glob, int(archive_costs[1]))
