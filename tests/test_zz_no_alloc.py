"""The allocation-free entry point (snn_select_device_noalloc) against
snn_select_device, through the C ABI: every result bit-identical, on two
calls in a row whose second problem is smaller (the fixed work array keeps
the first call's contents), in both forms and both directions; a problem
larger than the fixed array is rejected with an error, and the context stays
usable."""
import numpy as np
import pytest

from conftest import assert_same_result
from speedynn_amd import _lib, datasets as ds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    print("engine: creating context on device 0", flush=True)
    return _lib.Engine(0)


@pytest.fixture(scope="module")
def mem():
    return _lib.DeviceMemory()


def _compare(engine, mem, X, y, k_vals, loss, **kw):
    n, d = X.shape
    xp, yp = mem.upload(X), mem.upload(y)
    try:
        ref = engine.select(None, None, k_vals, [loss], device_inputs=(xp, yp, n, d),
                            candidate_errors=True, **kw)
        got = engine.select(None, None, k_vals, [loss], device_inputs=(xp, yp, n, d),
                            candidate_errors=True, noalloc=True, **kw)
    finally:
        mem.sync()
        mem.free(xp)
        mem.free(yp)
    assert_same_result(got, ref)
    steps = ref["steps"]
    assert got["steps"] == steps
    np.testing.assert_array_equal(got["candidate_errors"][:steps].view(np.uint32),
                                  ref["candidate_errors"][:steps].view(np.uint32))
    assert got["evals"] == ref["evals"]
    return got


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
def test_noalloc_matches_select_device_two_calls(engine, mem, form):
    # first call: forward classification, n = 1500, d = 40
    X1, y1, _ = ds.classification_argmax(301, 1500, 40, 4, 3, 0.05)
    print(f"noalloc form={form} call 1: forward n=1500 d=40", flush=True)
    _compare(engine, mem, X1, y1, [5, 10], 4, method=1, model_type=1, k_fold=1500,
             num_features=6, form=form)
    # second call, smaller: backward regression, n = 700, d = 24
    X2, y2, _ = ds.regression_sum(302, 700, 24, 3, 0.5)
    print(f"noalloc form={form} call 2: backward n=700 d=24", flush=True)
    _compare(engine, mem, X2, y2, [7], 0, method=0, model_type=0, k_fold=700,
             num_features=10, form=form)


def test_noalloc_rejects_problems_beyond_the_fixed_array(engine, mem):
    n, d = 5000, 8  # the distance matrix alone is 100 MB
    assert 4 * n * n > _lib.NOALLOC_SCRATCH_BYTES
    X, y, _ = ds.classification_sign(303, n, d, 2, 0.0)
    xp, yp = mem.upload(X), mem.upload(y)
    try:
        print("noalloc: oversized problem n=5000 d=8", flush=True)
        with pytest.raises(_lib.EngineError) as err:
            engine.select(None, None, [3], [4], method=1, model_type=1, k_fold=n,
                          num_features=2, form=1, device_inputs=(xp, yp, n, d), noalloc=True)
        assert err.value.code == _lib.SNN_EINVAL
        assert "fixed scratch" in str(err.value)
    finally:
        mem.sync()
        mem.free(xp)
        mem.free(yp)
    # the context and the fixed array still serve a problem that fits
    X2, y2, _ = ds.classification_sign(304, 400, 12, 3, 0.05)
    _compare(engine, mem, X2, y2, [5], 4, method=1, model_type=1, k_fold=400, num_features=4,
             form=1)
