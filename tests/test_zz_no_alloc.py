"""The allocation-free gradient entry point (faco_prior_gradients_noalloc)
against faco_prior_gradients through the C ABI: the same kernels over the
same partial rows, so the gradients and the loss must agree bit for bit, on
two calls in a row (the second with fewer edges, over the first call's
leftover scratch) and at the maximum size; sizes beyond the maximum are
rejected before any device work.
"""
import ctypes

import numpy as np
import pytest

from gplsaco_amd._native import check, lib
from tests.conftest import announce

pytestmark = pytest.mark.gpu

LIMIT = 8192


def _inputs(edges, seed):
    rng = np.random.default_rng(seed)
    features = rng.uniform(0.0, 4.0, size=(edges, 8)).astype(np.float32)
    labels = (rng.random(edges) < 0.1).astype(np.uint8)
    return features, labels


def _gradients(entry, prior, features, labels, pos_weight):
    P = lib().faco_prior_param_count()
    grads = np.zeros(P, dtype=np.float32)
    loss = ctypes.c_float()
    check(entry(prior, np.ascontiguousarray(features.reshape(-1)), labels, labels.size, pos_weight, grads,
                ctypes.byref(loss)))
    return grads, loss.value


@pytest.fixture(scope="module")
def prior():
    handle = ctypes.c_void_p()
    check(lib().faco_prior_create(1234, ctypes.byref(handle)))
    yield handle
    lib().faco_prior_destroy(handle)


@pytest.mark.parametrize("sizes", [(5000, 700), (LIMIT, 33), (64, 1)])
def test_noalloc_gradients_equal_the_allocating_entry_point(prior, sizes):
    for k, edges in enumerate(sizes):
        features, labels = _inputs(edges, 10 * edges + k)
        announce("noalloc-gradients", edges)
        want_g, want_loss = _gradients(lib().faco_prior_gradients, prior, features, labels, 2.0)
        got_g, got_loss = _gradients(lib().faco_prior_gradients_noalloc, prior, features, labels, 2.0)
        np.testing.assert_array_equal(got_g.view(np.uint32), want_g.view(np.uint32))
        assert np.float32(got_loss).view(np.uint32) == np.float32(want_loss).view(np.uint32)
        assert np.isfinite(got_g).all() and np.abs(got_g).sum() > 0


def test_noalloc_rejects_sizes_beyond_its_scratch(prior):
    features, labels = _inputs(LIMIT + 1, 3)
    with pytest.raises(ValueError, match="at most"):
        _gradients(lib().faco_prior_gradients_noalloc, prior, features, labels, 1.0)
    with pytest.raises(ValueError):
        _gradients(lib().faco_prior_gradients_noalloc, prior, features[:0], labels[:0], 1.0)
