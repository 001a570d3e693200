"""Single selections replayed from step graphs (params.use_graph): every
replay must equal the stream-issued selection bit for bit.  Graphs persist in
the context, so later calls replay what earlier calls captured; the calls are
interleaved with stream-issued selections and a different shape, and every
one is checked."""
import numpy as np
import pytest

from conftest import assert_same_result
from oracle import oracle as O
from speedynn_amd import _lib, datasets as ds

pytestmark = pytest.mark.gpu

CASES = [
    # seed, n, d, k_vals, k_fold, model, method, num_features
    (81, 600, 40, [5, 10], 600, "classification", "forward", 6),   # sorted step, walk, pruned
    (82, 512, 24, [7], 512, "regression", "backward", 4),          # M refreshes at d/2, d/4
    (83, 300, 70, [3, 8], 30, "classification", "forward", 4),     # k-fold, two blocks
    (84, 250, 12, [70], 250, "regression", "forward", 3),          # lists longer than 64
]


@pytest.fixture(scope="module")
def engine():
    print("engine: creating context on device 0", flush=True)
    return _lib.Engine(0)


def _data(seed, n, d, model):
    if model == "classification":
        X, y, _ = ds.classification_argmax(seed, n, d, 3, 3, 0.1)
    else:
        X, y, _ = ds.regression_sum(seed, n, d, 3, 0.5)
    return X, y


def _run(engine, case, form, graph):
    seed, n, d, k_vals, k_fold, model, method, nf = case
    X, y = _data(seed, n, d, model)
    code = 4 if model == "classification" else 0
    return engine.select(X, y, k_vals, [code], method=1 if method == "forward" else 0,
                         model_type=1 if model == "classification" else 0, k_fold=k_fold,
                         num_features=nf, form=form, candidate_errors=True, use_graph=graph)


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"seed{c[0]}")
def test_graph_replays_equal_stream(engine, case, form):
    seed, n, d, k_vals, k_fold, model, method, nf = case
    print(f"graph case seed={seed} n={n} d={d} form={form}", flush=True)
    X, y = _data(seed, n, d, model)
    loss = "zero_one" if model == "classification" else "rmse"
    o = O.select(X, y, k_vals, nf, k_fold, [loss], False, model, method, record_candidates=True)
    ref = _run(engine, case, form, False)
    assert_same_result(ref, o, "stream")
    other = CASES[(CASES.index(case) + 1) % len(CASES)]
    for rep in range(3):
        g = _run(engine, case, form, True)
        assert g["ms_search"] == 0.0 and g["ms_total"] > 0.0  # steps came from graphs
        assert_same_result(g, o, f"graph replay {rep}")
        steps = g["steps"]
        np.testing.assert_array_equal(g["candidate_errors"][:steps].view(np.uint32),
                                      ref["candidate_errors"][:steps].view(np.uint32))
        # eager work between replays: a stream-issued selection of another shape
        s = _run(engine, other, form, False)
        assert s["steps"] > 0
