"""GPU parity: the HIP engine against the reference's golden vectors and the
CPU oracle, through the C ABI.  Integer and index results must be bit-exact,
and so must every float error (both sides compute the same float32 ops in the
same order)."""
import base64
import json
import os

import numpy as np
import pytest

from conftest import (GOLDEN, assert_same_result, candidate_error_bits, expected_arrays,
                      golden_paths, load_case)
from oracle import oracle as O
from speedynn_amd import _lib, datasets as ds

pytestmark = pytest.mark.gpu

LOSS_CODES = {"rmse": 0, "rmse_norm": 1, "mad": 2, "mad_norm": 3, "zero_one": 4, "std": 5,
              "rmse_norm_mad_norm_std": 6, "rmse_norm_mad_norm_galaxies": 7, "rmse_std": 8}


@pytest.fixture(scope="module")
def engine():
    print("engine: creating context on device 0", flush=True)
    return _lib.Engine(0)


def run(engine, X, y, inp, form, candidate_errors=False):
    codes = [LOSS_CODES[l] for l in inp["loss_functions"]]
    return engine.select(X, y, inp["k_vals"], codes,
                         method=1 if inp["method"] == "forward" else 0,
                         model_type=1 if inp["model_type"] == "classification" else 0,
                         k_fold=inp["k_fold"], num_features=inp["num_features"],
                         auto_feature=inp["auto_feature"], form=form,
                         candidate_errors=candidate_errors)


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
@pytest.mark.parametrize("path", golden_paths(), ids=os.path.basename)
def test_engine_matches_reference_golden(engine, path, form):
    rec, X, y = load_case(path)
    inp = rec["inputs"]
    print(f"case {os.path.basename(path)} n={X.shape[0]} d={X.shape[1]} form={form}", flush=True)
    got = run(engine, X, y, inp, form)
    assert_same_result(got, expected_arrays(rec), os.path.basename(path))


RANDOM = [
    # seed, n, d, k_vals, k_fold, model, method, num_features
    (61, 300, 12, [1, 3, 5], 300, "classification", "forward", 5),
    (62, 257, 9, [2, 4, 8], 13, "classification", "backward", 4),
    (63, 400, 10, [5, 10, 20], 400, "regression", "forward", 4),
    (64, 333, 8, [15], 333, "regression", "backward", 3),
    (65, 130, 70, [7, 3], 130, "classification", "forward", 3),
    (66, 190, 6, [32, 50], 190, "regression", "forward", 3),
    # longest lists backward (the 64-entry list, and the 50-entry list of k = 33 over two
    # candidate blocks and two folds), and k = 20 forward over two candidate blocks
    (67, 260, 20, [64], 260, "classification", "backward", 15),
    (68, 300, 70, [24, 33], 150, "regression", "backward", 66),
    (69, 280, 90, [20], 280, "classification", "forward", 4),
]


def _random_case(seed, n, d, model):
    if model == "classification":
        X, y, _ = ds.classification_argmax(seed, n, d, 3, 3, 0.1)
        if seed % 2:
            X = ds.quantized(X, 0.25)
    else:
        X, y, _ = ds.regression_sum(seed, n, d, 3, 0.5)
    return X, y


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
@pytest.mark.parametrize("case", RANDOM, ids=lambda c: f"seed{c[0]}")
def test_every_candidate_error_matches_oracle(engine, case, form):
    seed, n, d, k_vals, k_fold, model, method, nf = case
    X, y = _random_case(seed, n, d, model)
    loss = "zero_one" if model == "classification" else "rmse"
    print(f"random case seed={seed} n={n} d={d} form={form}", flush=True)
    o = O.select(X, y, k_vals, nf, k_fold, [loss], False, model, method, record_candidates=True)
    inp = dict(k_vals=k_vals, loss_functions=[loss], method=method, model_type=model,
               k_fold=k_fold, num_features=nf, auto_feature=False)
    g = run(engine, X, y, inp, form, candidate_errors=True)
    assert_same_result(g, o)
    steps = g["steps"]
    np.testing.assert_array_equal(g["candidate_errors"][:steps].view(np.uint32),
                                  o["candidate_errors"][:steps].view(np.uint32))


LONG = [
    # lists longer than 64 (global-memory lists): forward LOO, backward over uneven folds,
    # and k above the eligible points (unfilled slots keep index 0); two candidate blocks
    (71, 300, 10, [70, 100, 3], 300, "classification", "forward", 4),
    (72, 257, 9, [65, 3], 13, "regression", "backward", 4),
    (73, 90, 70, [100, 5], 90, "classification", "forward", 2),
]


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
@pytest.mark.parametrize("case", LONG, ids=lambda c: f"seed{c[0]}")
def test_lists_longer_than_64_match_oracle(engine, case, form):
    """k values above 64 (the reference takes any k, cpu.c:50-51): every
    candidate's error at every step against the oracle, bit for bit."""
    seed, n, d, k_vals, k_fold, model, method, nf = case
    X, y = _random_case(seed, n, d, model)
    loss = "zero_one" if model == "classification" else "rmse"
    print(f"long-list case seed={seed} n={n} d={d} k={k_vals} form={form}", flush=True)
    o = O.select(X, y, k_vals, nf, k_fold, [loss], False, model, method, record_candidates=True)
    inp = dict(k_vals=k_vals, loss_functions=[loss], method=method, model_type=model,
               k_fold=k_fold, num_features=nf, auto_feature=False)
    g = run(engine, X, y, inp, form, candidate_errors=True)
    assert_same_result(g, o)
    steps = g["steps"]
    np.testing.assert_array_equal(g["candidate_errors"][:steps].view(np.uint32),
                                  o["candidate_errors"][:steps].view(np.uint32))


NONFINITE = [
    # seed, n, d, k_vals, k_fold, model, method, num_features
    (63, 400, 10, [5, 10, 20], 400, "regression", "forward", 4),
    (62, 257, 9, [2, 4, 8], 13, "classification", "backward", 4),
]


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
@pytest.mark.parametrize("case", NONFINITE, ids=lambda c: f"seed{c[0]}-{c[6]}")
def test_nonfinite_features_match_oracle(engine, case, form):
    """X holding NaN in two columns and +-inf in a third: a distance through a
    NaN or an infinity is NaN or inf, which never passes the strict '<' of the
    list (cpu.c:149-165), so such points stay out of the lists, and the
    candidates whose errors come out NaN are never chosen.  The precision
    form takes its non-finite paths here (no sorted first step, the walk's
    fallback, NaN slot keys).  Errors are compared bit for bit, NaN with NaN
    (the payload of a NaN is not part of the result)."""
    seed, n, d, k_vals, k_fold, model, method, nf = case
    X, y = _random_case(seed, n, d, model)
    X = X.copy()
    X[5, 2] = np.nan
    X[17, 7] = np.nan
    X[33, 4], X[34, 4], X[40, 4] = np.inf, -np.inf, np.inf
    loss = "zero_one" if model == "classification" else "rmse"
    print(f"nonfinite case seed={seed} n={n} d={d} form={form}", flush=True)
    o = O.select(X, y, k_vals, nf, k_fold, [loss], False, model, method, record_candidates=True)
    inp = dict(k_vals=k_vals, loss_functions=[loss], method=method, model_type=model,
               k_fold=k_fold, num_features=nf, auto_feature=False)
    g = run(engine, X, y, inp, form, candidate_errors=True)
    assert_same_result(g, o)
    steps = g["steps"]
    got, exp = g["candidate_errors"][:steps], o["candidate_errors"][:steps]
    assert np.isnan(exp).any()  # the case does reach NaN errors
    np.testing.assert_array_equal(np.isnan(got), np.isnan(exp))
    fin = ~np.isnan(exp)
    np.testing.assert_array_equal(got[fin].view(np.uint32), exp[fin].view(np.uint32))


@pytest.mark.parametrize("method", ["forward", "backward"])
def test_distance_rows_bit_exact(engine, method):
    X, _, _ = ds.regression_sum(70, 500, 40, 3, 0.5)
    order = [7, 3, 33, 0, 19]
    rows = [0, 1, 250, 499]
    got, _ = engine.distance_rows(X, 1 if method == "forward" else 0, order, rows, form=0)
    for r, q in enumerate(rows):
        np.testing.assert_array_equal(got[r].view(np.uint32),
                                      O.distance_row(X, q, method, order).view(np.uint32))


def test_compensated_rows_keep_plain_bits_and_gain_accuracy(engine):
    X, _, _ = ds.regression_sum(71, 400, 64, 3, 0.5)
    X = X * np.float32(37.0)
    order = list(range(0, 64, 2))
    rows = [3, 200]
    plain, _ = engine.distance_rows(X, 1, order, rows, form=0)
    hi, lo = engine.distance_rows(X, 1, order, rows, form=1)
    np.testing.assert_array_equal(plain.view(np.uint32), hi.view(np.uint32))
    for r, q in enumerate(rows):
        exact = O.distance_row_f64(X, q, order)
        nz = exact > 0
        e_plain = np.abs(plain[r][nz] - exact[nz]) / exact[nz]
        e_comp = np.abs(hi[r][nz].astype(np.float64) + lo[r][nz] - exact[nz]) / exact[nz]
        assert e_comp.max() < 1e-9
        assert e_comp.max() < e_plain.max()


def test_backward_full_matrix_relative_error(engine):
    X, _, _ = ds.regression_sum(72, 300, 256, 3, 0.5)
    got, lo = engine.distance_rows(X, 0, [], [0, 150, 299], form=1)
    for r, q in enumerate([0, 150, 299]):
        exact = O.distance_row_f64(X, q, list(range(256)))
        nz = exact > 0
        rel = np.abs(got[r][nz] - exact[nz]) / exact[nz]
        assert rel.max() < 1e-5
        rel2 = np.abs(got[r][nz].astype(np.float64) + lo[r][nz] - exact[nz]) / exact[nz]
        assert rel2.max() < 1e-11


def test_losses_match_reference_known_answers(engine):
    with open(os.path.join(GOLDEN, "losses_votes.json")) as fh:
        rec = json.load(fh)
    sets = {k: [np.frombuffer(base64.b64decode(s), np.float32) for s in v]
            for k, v in rec["sets"].items()}
    for item in rec["losses"]:
        y, p = sets[item["set"]]
        got = np.float32(engine.loss(y, p, LOSS_CODES[item["loss"]])).view(np.uint32)
        assert int(got) == item["value_bits"], item


def test_dropin_api_and_swig_surface_on_c0():
    from speedynn_amd import IncrementalNNFeatureSelection
    from speedynn_amd import nearest_neighbors_feature_selection_gpu as M
    rec, X, y = load_case(os.path.join(GOLDEN, "c0.json"))
    exp = expected_arrays(rec)
    m = IncrementalNNFeatureSelection(X, y, k_vals=[10], num_features=10, k_fold=2000,
                                      method="forward", loss_functions=["zero_one"],
                                      model_type="classification")
    m.train()
    np.testing.assert_array_equal(m.get_selected_ordering(), exp["selected_ordering"])
    np.testing.assert_array_equal(m.get_selected_errors()[:, 0].view(np.uint32),
                                  exp["selected_errors"].view(np.uint32))
    nf = M.train_forward_selection(X, y, np.array([10], np.int32), 10, 2000,
                                   np.array([4], np.int32), 0, 1)
    assert nf == exp["num_final"]
    order = np.zeros(50, np.int32)
    M.get_selected_ordering(order)
    np.testing.assert_array_equal(order, exp["selected_ordering"])
    assert M.get_final_optimal_model_k() == exp["optimal_k"]


def test_invalid_arguments_return_errors(engine):
    X, y, _ = ds.classification_sign(80, 50, 4, 2, 0.0)
    with pytest.raises(_lib.EngineError):
        engine.select(X, y, [3], [4], method=1, model_type=1, k_fold=51, num_features=2)
    with pytest.raises(_lib.EngineError):
        engine.select(X, y, [0], [4], method=1, model_type=1, k_fold=50, num_features=2)
    with pytest.raises(_lib.EngineError):
        engine.select(X, y, [3], [9], method=1, model_type=1, k_fold=50, num_features=2)
    with pytest.raises(_lib.EngineError):
        engine.select(X, y, [3], [4], method=1, model_type=1, k_fold=50, num_features=5)
    # the context stays usable after rejected calls
    r = engine.select(X, y, [3], [4], method=1, model_type=1, k_fold=50, num_features=2)
    assert r["num_final"] == 2


REDUCED = sorted(os.path.join(GOLDEN, f) for f in os.listdir(GOLDEN)
                 if f.endswith("_oracle.json") and "_n" in f)


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
@pytest.mark.parametrize("path", REDUCED, ids=os.path.basename)
def test_workload_recipes_at_reduced_size_match_oracle(engine, path, form):
    """The n=65536 forward workload's recipe (log-uniform feature scales, 32
    informative features, k = 20: two candidate blocks' worth of slots plus a
    partly filled last block) at n = 3000 with all 512 features, and the
    n=32768 backward regression recipe (noisy column copies) at n = 1000,
    d = 320, down to 16: selections, per-step errors and every candidate's
    error at every step against the oracle."""
    rec, X, y = load_case(path)
    print(f"{rec['name']} form={form}", flush=True)
    got = run(engine, X, y, rec["inputs"], form, candidate_errors=True)
    assert_same_result(got, expected_arrays(rec), rec["name"])
    steps = got["steps"]
    np.testing.assert_array_equal(got["candidate_errors"][:steps].view(np.uint32),
                                  candidate_error_bits(rec))


C1 = os.path.join(GOLDEN, "c1_oracle.json")


@pytest.mark.slow
@pytest.mark.skipif(not os.path.exists(C1), reason="c1 oracle fixture not generated")
@pytest.mark.parametrize("form", [0, 1], ids=["plain", "precision"])
def test_c1_full_size_matches_oracle(engine, form):
    """n=20000, d=200, k=10, 3 classes, forward to 20: the selected sequence, the
    per-step errors (integer LOO counts / n) and all 3810 candidate errors."""
    rec, X, y = load_case(C1)
    print(f"c1 form={form}", flush=True)
    got = run(engine, X, y, rec["inputs"], form, candidate_errors=True)
    assert_same_result(got, expected_arrays(rec), "c1")
    steps = got["steps"]
    exp_c = np.array(rec["expected"]["candidate_errors_bits"], np.uint32).reshape(steps, -1)
    np.testing.assert_array_equal(got["candidate_errors"][:steps].view(np.uint32), exp_c)


@pytest.mark.parametrize("name", ["reg_bwd_loo.json", "tiny_cls_fwd_loo.json"])
def test_compensated_precision_selects_the_same(engine, name):
    """Keeping D's compensation term (params.compensated) changes no bit of
    the selection: the search reads D in the reference's bits either way."""
    rec, X, y = load_case(os.path.join(GOLDEN, name))
    inp = rec["inputs"]
    codes = [LOSS_CODES[l] for l in inp["loss_functions"]]
    print(f"compensated precision form on {name}", flush=True)
    got = engine.select(X, y, inp["k_vals"], codes,
                        method=1 if inp["method"] == "forward" else 0,
                        model_type=1 if inp["model_type"] == "classification" else 0,
                        k_fold=inp["k_fold"], num_features=inp["num_features"],
                        auto_feature=inp["auto_feature"], form=1, compensated=True)
    assert_same_result(got, expected_arrays(rec), name)


@pytest.mark.parametrize("n,d", [(1500, 300), (1001, 37), (8300, 37)])
def test_mfma_gram_within_1e5_of_fp64(engine, n, d):
    """(a) on the matrix cores: the BF16x3 MFMA Gram (SNN_FORM_MFMA) against
    the FP64 oracle, on unit-scale Gaussian features and on features with
    scales 10^U(-2,2).  Tolerance 1e-5 relative; the diagonal is exactly 0 and
    the matrix exactly symmetric.  No n is a multiple of its tile (128 points
    below n = 8192, 256 from there on: partial tiles); n = 1001 is odd, so
    every mirror store takes the scalar path, and d = 37 pads the last
    16-feature step with zeros.  n = 8300 runs the 16-wave 256-point tiles,
    whose diagonal tiles hold waves below, on and above the diagonal."""
    Z = ds.gaussian_matrix(91, n, d)
    scales = ds.exp2(np.log2(10.0) * (4.0 * ds.uniforms(92, 0, d) - 2.0))
    cases = {"unit": Z, "scaled": (Z.astype(np.float64) * scales[None, :]).astype(np.float32)}
    rows = sorted({0, 7, 70, 200, 777, n // 2 + 3, n - 1})
    for name, X in cases.items():
        print(f"mfma gram {name} n={n} d={d}", flush=True)
        got, lo = engine.distance_rows(X, 0, [], rows, form=2)
        assert lo is None
        worst = 0.0
        for r, q in enumerate(rows):
            ref = O.distance_row_f64(X, q, list(range(d)))
            off = np.arange(n) != q
            rel = np.abs(got[r][off].astype(np.float64) - ref[off]) / ref[off]
            worst = max(worst, float(rel.max()))
            assert got[r][q] == 0.0
        # one triangle is computed and mirrored: D is exactly symmetric
        for r, q in enumerate(rows):
            for r2, q2 in enumerate(rows):
                assert got[r][q2] == got[r2][q], (name, q, q2)
        print(f"mfma gram {name}: max relative error {worst:.3e}", flush=True)
        assert worst <= 1e-5, (name, worst)


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_c2_full_size_precision_equals_plain(engine):
    """n=65536, d=512, k=20, forward to 32 (beyond the oracle's reach at this
    size; its recipe is checked against the oracle at n=3000 above): the
    precision form (sorted first step, warm bounds, row lists, exact pruning)
    and the plain form (dense search, the reference's operation order) give
    the same selected sequence, per-step errors and every candidate's error
    at every step, bit for bit."""
    w = ds.config("c2")
    X, y = w["X"], w["y"]
    n, d = X.shape
    mem = _lib.DeviceMemory()
    xp, yp = mem.upload(X), mem.upload(y)
    try:
        got = {}
        for form in (1, 0):
            print(f"c2 full size form={form}", flush=True)
            got[form] = engine.select(None, None, w["k_vals"], [4], method=1, model_type=1,
                                      k_fold=n, num_features=w["num_features"], form=form,
                                      candidate_errors=True, device_inputs=(xp, yp, n, d))
    finally:
        mem.sync()
        mem.free(xp)
        mem.free(yp)
    assert_same_result(got[1], got[0], "c2")
    steps = got[0]["steps"]
    assert steps == 32
    np.testing.assert_array_equal(got[1]["candidate_errors"][:steps].view(np.uint32),
                                  got[0]["candidate_errors"][:steps].view(np.uint32))
