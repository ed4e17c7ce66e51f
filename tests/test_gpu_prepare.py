"""Device-side problem preparation (register_problems_device) against the
host preparation (register_problem, itself pinned to the oracle and the
reference's checklist_two_opt vectors): candidate rows, initial tour and
cost must be identical, and so must every search that starts from them.

Cases: uniform and clustered instances, instances with many repeated
points (distance ties and zero distances), the smallest sizes (3..10
cities, where the candidate widths clamp), both sides of the 1024-city
boundary where the ant kernel's tour switches from linked lists to arrays,
and a mixed-size engine with sizes 200..2000 in one call.
"""
import numpy as np
import pytest

from gplsaco_amd.instances import clustered_instance, uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def crowded_instance(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 12, size=(n, 2)).astype(np.float64)   # many equal points and equal distances


def _same_data(host, device, key):
    a, b = host.problem_data(key), device.problem_data(key)
    for name in ("primary", "backup", "ls", "initial_tour"):
        np.testing.assert_array_equal(a[name], b[name], err_msg=f"{name} of instance {key}")
    assert a["initial_cost"] == b["initial_cost"], key


def _pair(ext, engine_type, n, colonies, xys, keys, **settings):
    s = ext.FixedFacoSettings()
    for k, v in settings.items():
        setattr(s, k, v)
    host, device = engine_type(n, colonies, s), engine_type(n, colonies, s)
    for k, xy in zip(keys, xys):
        host.register_problem(k, xy)
    out = device.register_problems_device(keys, xys)
    assert out["instances"] == len(keys) and out["seconds"] > 0
    return host, device


@pytest.mark.parametrize("n,kind", [(200, "uniform"), (500, "clustered"), (300, "crowded"), (1024, "uniform"),
                                    (1025, "clustered"), (3, "uniform"), (5, "crowded"), (9, "uniform"),
                                    (10, "crowded")])
def test_device_preparation_equals_host(ext, n, kind):
    make = {"uniform": lambda i: uniform_instance(n, seed=31, instance=i),
            "clustered": lambda i: clustered_instance(n, seed=31, instance=i),
            "crowded": lambda i: crowded_instance(n, seed=i)}[kind]
    xys = [make(i) for i in range(3)]
    keys = [500 + i for i in range(3)]
    announce("prepare", n, kind)
    host, device = _pair(ext, ext.FacoBatchEngine, n, 6, xys, keys, ants=32 if n > 10 else 8)
    for k in keys:
        _same_data(host, device, k)
    # every search from the prepared data is the same (heuristic, local
    # scales and the distance table included)
    task_keys = np.array([keys[c % 3] for c in range(6)], dtype=np.uint64)
    seeds = np.arange(6, dtype=np.uint64)
    ants = host.settings.ants
    a = host.evaluate_faco_evaluations(task_keys, seeds, ants * 3)
    b = device.evaluate_faco_evaluations(task_keys, seeds, ants * 3)
    assert [i["tour"] for i in a["items"]] == [i["tour"] for i in b["items"]]
    assert a["completed_ls_evaluations"] == b["completed_ls_evaluations"]
    if n >= 64:
        policy = {"policy_spec_id": 1, "kind": "rule", "mne_level": 0, "max_mne_level": 3, "region": 1,
                  "restart_mode": "none", "restart_period": 0, "restart_probability": 0.0, "stagnation_step": 2,
                  "restart_stagnation": 5, "restart_cooldown": 3}
        x = host.evaluate_baseline_evaluations(task_keys, seeds, ants * 4, policy, decision_iterations=[4])
        y = device.evaluate_baseline_evaluations(task_keys, seeds, ants * 4, policy, decision_iterations=[4])
        assert [i["tour"] for i in x["items"]] == [i["tour"] for i in y["items"]]
        assert x["decisions"] == y["decisions"]


def test_mixed_sizes_in_one_call_and_repeated_keys(ext):
    rng = np.random.default_rng(64)
    sizes = [int(v) for v in rng.integers(200, 2001, size=24)]
    xys = [(clustered_instance if i % 2 else uniform_instance)(m, seed=64, instance=i) for i, m in enumerate(sizes)]
    keys = [9000 + i for i in range(len(sizes))]
    announce("prepare-mixed", len(sizes), max(sizes))
    host, device = _pair(ext, ext.MixedFacoBatchEngine, max(sizes), len(sizes), xys, keys)
    for k in keys:
        _same_data(host, device, k)
    again = device.register_problems_device(keys[:2], xys[:2])      # registered already: skipped
    assert again["instances"] == 2
    changed = xys[0].copy()
    changed[0, 0] = 1.0 if changed[0, 0] != 1.0 else 2.0
    with pytest.raises(ValueError):
        device.register_problems_device([keys[0]], [changed])
    with pytest.raises(ValueError):
        device.register_problems_device([77, 77], [xys[0], changed])   # one key, two coordinate sets
    print(f"prepared {len(sizes)} instances n in [200, 2000]: device {device.register_problems_device([1], [xys[1]])}",
          flush=True)
