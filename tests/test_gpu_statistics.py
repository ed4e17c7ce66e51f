"""Stochastic parity: the device's final tour lengths over many seeds against
the same algorithm driven by an unrelated random generator.

The device is bit-exact with the oracle under the shared Philox streams
(every other GPU test), so its tour-length distribution is the oracle's.
The reference draws its uniforms from cuRAND instead, which is not available
here; what remains to show is that the choice of uniform source does not
move the distribution. The oracle's alternate stream (SplitMix64, unrelated
to Philox) stands in for any other generator:

- the CPU-runnable configuration (one uniform TSP100, 20-nearest-neighbour
  candidates, 32 ants, 1000 iterations, 2-opt + Or-opt): the device's mean
  over seeds 0..29 lies in the 95% confidence interval (Student t) of the
  alternate stream's mean over the same seeds;
- the TSP1000 headline shape (128 ants, 100 iterations, 2-opt): two samples
  of 160 seeds each pass Welch's t-test and the two-sample Kolmogorov-Smirnov
  test at the 1% level. (A 30-seed CI check is not used here: two
  independent 30-sample means of the same distribution fall outside each
  other's 95% CI about one time in six.)
"""
import os

import numpy as np
import pytest
from scipy import stats

from gplsaco_amd.instances import uniform_instance
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def _device_costs(ext, xy, seeds, ants, iterations, widths=None, or_opt=False):
    s = ext.FixedFacoSettings()
    s.ants = ants
    if widths:
        s.primary_width, s.ls_width = widths
    engine = ext.FacoBatchEngine(xy.shape[0], seeds.size, s)
    engine.register_problem(1, xy)
    engine.set_local_search("2opt+oropt" if or_opt else "2opt")
    out = engine.evaluate_faco_evaluations(np.ones(seeds.size, dtype=np.uint64), seeds, ants * iterations)
    return np.array([item["cost"] for item in out["items"]])


def _alternate_costs(oracle_mod, xy, seeds, ants, iterations, widths=None, or_opt=False):
    O = oracle_mod
    kw = {"primary_width": widths[0], "ls_width": widths[1]} if widths else {}
    with O.alternate_stream():
        out = O.run_faco(xy[None].astype(np.int32), [1], seeds, O.default_settings(ants=ants, **kw), iterations, 8,
                         threads=min(16, os.cpu_count() or 1), or_opt=or_opt)
    return out["costs"].astype(float)


def test_tsp100_mean_inside_the_alternate_streams_confidence_interval(ext, oracle_mod):
    xy = uniform_instance(100, seed=0)
    seeds = np.arange(30, dtype=np.uint64)
    announce("statistics", "TSP100", 30)
    device = _device_costs(ext, xy, seeds, 32, 1000, widths=(20, 20), or_opt=True)
    other = _alternate_costs(oracle_mod, xy, seeds, 32, 1000, widths=(20, 20), or_opt=True)
    half = stats.t.ppf(0.975, other.size - 1) * other.std(ddof=1) / np.sqrt(other.size)
    print(f"TSP100 device mean {device.mean():.1f} alternate {other.mean():.1f} +- {half:.1f}", flush=True)
    assert abs(device.mean() - other.mean()) <= half


def test_tsp1000_distribution_matches_the_alternate_stream(ext, oracle_mod):
    xy = uniform_instance(1000, seed=0)
    seeds = np.arange(1000, 1160, dtype=np.uint64)
    announce("statistics", "TSP1000", seeds.size)
    device = _device_costs(ext, xy, seeds, 128, 100)
    other = _alternate_costs(oracle_mod, xy, seeds, 128, 100)
    welch = stats.ttest_ind(device, other, equal_var=False).pvalue
    ks = stats.ks_2samp(device, other).pvalue
    print(f"TSP1000 device mean {device.mean():.1f} sd {device.std(ddof=1):.1f}; alternate {other.mean():.1f} "
          f"sd {other.std(ddof=1):.1f}; Welch p {welch:.3f}, KS p {ks:.3f}", flush=True)
    assert welch > 0.01 and ks > 0.01
