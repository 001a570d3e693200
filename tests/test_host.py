"""CPU-side checks: the C ABI library loads and exports every declared symbol,
the Python API's host logic matches the reference's, and the generator is
deterministic.  No GPU compute is called here."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from speedynn_amd import _lib, datasets as ds
from speedynn_amd.base import IncrementalNNFeatureSelection

HEADER = os.path.join(ROOT, "include", "speedynn.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*([a-z_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("engine library not built (run python -m speedynn_amd.build)")
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()


class FakeEngine:
    """Returns a canned result so the host-side bookkeeping can be checked."""

    def __init__(self, res):
        self.res = res
        self.calls = []

    def select(self, X, y, k_vals, loss_codes, **kw):
        self.calls.append((k_vals, loss_codes, kw))
        return self.res


def _fake(monkeypatch, res):
    eng = FakeEngine(res)
    monkeypatch.setattr(_lib, "default_engine", lambda: eng)
    return eng


def test_constructor_defaults_and_coercion(monkeypatch):
    X = np.arange(12, dtype=np.float64).reshape(4, 3)
    m = IncrementalNNFeatureSelection(X, np.array([1, 0, 1, 0]))
    assert m.k_vals.tolist() == list(range(1, 10)) and m.k_vals.dtype == np.int32
    assert m.Xtrain.dtype == np.float32 and m.Ytrain.dtype == np.float32
    assert m.method == "backward" and m.k_fold == 10 and m.loss_functions.tolist() == [0]
    m2 = IncrementalNNFeatureSelection(X, np.ones(4), auto_feature=True, num_features=1)
    assert m2.num_desired_features == 3
    with pytest.raises(AssertionError):
        IncrementalNNFeatureSelection(X, np.ones(4), num_features=4)
    with pytest.raises(AssertionError):
        IncrementalNNFeatureSelection(X, np.ones(4), num_features=0)


def test_forward_ranks_and_error_layout(monkeypatch):
    d = 5
    res = dict(num_final=3, optimal_k=7, steps=3,
               selected_dimensions=np.array([1, 0, 1, 0, 1], np.int32),
               selected_ordering=np.array([4, 0, 2, 0, 0], np.int32),
               selected_errors=np.arange(2 * d, dtype=np.float32))
    eng = _fake(monkeypatch, res)
    m = IncrementalNNFeatureSelection(np.zeros((6, d)), np.zeros(6), k_vals=[3, 7],
                                      num_features=3, method="forward",
                                      loss_functions=["zero_one", "rmse"],
                                      model_type="classification")
    m.train()
    kw = eng.calls[0][2]
    assert kw["method"] == 1 and kw["model_type"] == 1 and kw["num_features"] == 3
    assert eng.calls[0][1].tolist() == [4, 0]
    # speedynn/base.py:79-84: errors[l*dim:(l+1)*dim] becomes column l
    np.testing.assert_array_equal(m.get_selected_errors()[:, 0], np.arange(5))
    np.testing.assert_array_equal(m.get_selected_errors()[:, 1], np.arange(5, 10))
    # speedynn/base.py:99-103
    np.testing.assert_array_equal(m.get_feature_ranks(), [2, 0, 1, 0, 3])
    assert m.get_optimal_model_k() == 7


def test_backward_ranks(monkeypatch):
    d = 5
    res = dict(num_final=2, optimal_k=1, steps=3,
               selected_dimensions=np.array([0, 1, 0, 1, 0], np.int32),
               selected_ordering=np.array([0, 0, 4, 0, 2], np.int32),
               selected_errors=np.zeros(d, np.float32))
    _fake(monkeypatch, res)
    m = IncrementalNNFeatureSelection(np.zeros((6, d)), np.zeros(6), num_features=2)
    m.train()
    # speedynn/base.py:94-98: ranks[ordering[dim-i-1]] = -(dim-final) + i; ordering[4] is
    # the first feature removed (rank -3), ordering[2] the last (rank -1)
    np.testing.assert_array_equal(m.get_feature_ranks(), [-2, 0, -3, 0, -1])


def test_unknown_method_raises(monkeypatch):
    _fake(monkeypatch, {})
    m = IncrementalNNFeatureSelection(np.zeros((4, 2)), np.zeros(4), num_features=1,
                                      method="sideways")
    with pytest.raises(ValueError):
        m.train()


def test_swig_module_rejects_wrong_dtypes():
    from speedynn_amd import nearest_neighbors_feature_selection_gpu as M
    with pytest.raises(TypeError):
        M.train_forward_selection(np.zeros((4, 2)), np.zeros(4, np.float32),
                                  np.array([1], np.int32), 1, 4, np.array([4], np.int32), 0, 1)


def test_generator_is_deterministic_and_gaussian():
    a = ds.normals(3, 0, 200000)
    b = ds.normals(3, 0, 200000)
    np.testing.assert_array_equal(a, b)
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1) < 0.01
    assert abs(((a - a.mean()) ** 4).mean() / a.var() ** 2 - 3) < 0.05
    c = ds.choice(3, 1, 50, 50)
    assert sorted(c.tolist()) == list(range(50))
    assert ds.checksum(ds.normals(3, 0, 100)) != ds.checksum(ds.normals(4, 0, 100))


def test_workload_shapes():
    shapes = ds.batch_problem_shapes()
    assert len(shapes) == 64
    assert all(1000 <= n <= 16000 and 32 <= d <= 256 and k in (5, 10, 20) for n, d, k in shapes)
    w = ds.config("c0")
    assert w["X"].shape == (2000, 50) and set(np.unique(w["y"])) == {-1.0, 1.0}
    assert int((w["y"] == 1).sum()) + int((w["y"] == -1).sum()) == 2000


def test_reference_qsort_order_matches_the_c_library(tmp_path):
    """The device's median sort for NaN-holding arrays (qsort_order.h) leaves
    every array exactly as the C library's qsort does with the reference's
    comparator, NaN, infinities, signed zeros and ties included; so does its
    64-worker split used by one wave."""
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path / "qsort_order_check"
    src = os.path.join(ROOT, "tests", "qsort_order_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe), src],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    so, before any device call."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_reads_committed_traffic_summaries():
    """Every workload's bench line carries roofline.traffic from the committed
    rocprofv3 counter summary (profiles/traffic_search_<workload>.md)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for w in ("c1", "c2", "c3", "c4"):
        t, src = bench.traffic_per_launch(w)
        assert t is not None and t > 0, w
        assert src.startswith(f"profiles/traffic_search_{w}.md"), src
