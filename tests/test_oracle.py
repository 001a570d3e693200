"""Pins of the CPU oracle (oracle/oracle.c), independent of the GPU engine.

* every golden vector written by the reference's own compiled CPU module;
* a brute-force restatement on tiny inputs: sort all eligible points by
  (distance, index) with numpy float32 arithmetic, take the first k;
* the reference's get_loss_all / get_mode_array known answers;
* closed forms of the losses and the vote's tie rule;
* planted mistakes (wrong tie rule, dropped fold exclusion) fail the checks.
"""
import base64
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_same_result, expected_arrays, golden_paths, load_case
from oracle import oracle as O
from speedynn_amd import datasets as ds

MAXF = np.float32(3.402823466e38)


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


@pytest.mark.parametrize("path", golden_paths(), ids=os.path.basename)
def test_oracle_matches_reference_golden(path):
    rec, X, y = load_case(path)
    inp = rec["inputs"]
    got = O.select(X, y, inp["k_vals"], inp["num_features"], inp["k_fold"], inp["loss_functions"],
                   inp["auto_feature"], inp["model_type"], inp["method"])
    assert_same_result(got, expected_arrays(rec), os.path.basename(path))


def test_oracle_thread_count_invariant():
    rec, X, y = load_case(os.path.join(GOLDEN, "cls3_fwd_even_k.json"))
    inp = rec["inputs"]
    args = (X, y, inp["k_vals"], inp["num_features"], inp["k_fold"], inp["loss_functions"],
            inp["auto_feature"], inp["model_type"], inp["method"])
    a = O.select(*args, threads=1)
    b = O.select(*args, threads=4)
    assert_same_result(a, b)


# ------------------------------------------------------------ brute force

def brute_row(X, q, order, method):
    """D row in reference order with numpy float32 scalars."""
    n, d = X.shape
    X = X.astype(np.float32)
    if method == "forward":
        row = np.zeros(n, np.float32)
        for f in order:
            t = (X[q, f] - X[:, f]).astype(np.float32)
            row = (row + (t * t).astype(np.float32)).astype(np.float32)
    else:
        row = np.zeros(n, np.float32)
        for f in range(d - 1, -1, -1):
            t = (X[q, f] - X[:, f]).astype(np.float32)
            row = (row + (t * t).astype(np.float32)).astype(np.float32)
        for f in order:
            t = (X[q, f] - X[:, f]).astype(np.float32)
            row = (row - (t * t).astype(np.float32)).astype(np.float32)
    row[q] = 0.0
    return row


def brute_mode(labels):
    first, count = {}, {}
    for i, v in enumerate(labels):
        key = float(v)
        if key != key:  # NaN never equals anything: its own class
            key = ("nan", i)
        if key not in first:
            first[key] = i
            count[key] = 0
        count[key] += 1
    best = max(count.values())
    pos = min(first[k] for k in count if count[k] == best)
    return np.float32(labels[pos])


def brute_preds(X, y, order, method, f, k_vals, k_fold, model_type, tie_rule="index",
                drop_fold=False):
    n, _ = X.shape
    npf = n // k_fold
    mult = np.float32(1.0 if method == "forward" else -1.0)
    kmax = max(k_vals)
    preds = np.zeros((len(k_vals), n), np.float32)
    for q in range(n):
        row = brute_row(X, q, order, method)
        t = (X[q, f] - X[:, f]).astype(np.float32)
        dd = (row + mult * (t * t).astype(np.float32)).astype(np.float32)
        idx = np.arange(n)
        ok = (idx // npf) != (q // npf)
        if drop_fold:
            ok = idx != q
        ok &= dd < MAXF
        cand = idx[ok]
        key2 = cand if tie_rule == "index" else -cand
        srt = cand[np.lexsort((key2, dd[ok]))][:kmax]
        nn = np.concatenate([srt, np.zeros(kmax - len(srt), np.int64)])
        for m, k in enumerate(k_vals):
            lab = y[nn[:k]]
            if model_type == "regression":
                s = np.float32(0.0)
                for v in lab:
                    s = np.float32(s + v)
                preds[m, q] = np.float32(s / np.float32(k))
            else:
                preds[m, q] = brute_mode(lab)
    return preds


TINY = [
    # seed, n, d, k_vals, k_fold, model, method, order, f
    (31, 23, 5, [1, 3, 4], 23, "classification", "forward", [], 2),
    (32, 29, 6, [2, 5], 29, "classification", "forward", [4, 1], 3),
    (33, 30, 6, [3, 7], 4, "regression", "forward", [0], 5),
    (34, 25, 5, [2, 4], 25, "regression", "backward", [3], 1),
    (35, 31, 4, [6], 3, "classification", "backward", [], 0),
]


def _tiny_data(seed, n, d, model):
    if model == "classification":
        X, y, _ = ds.classification_argmax(seed, n, d, 2, 3, 0.1)
        X = ds.quantized(X, 0.5)   # many exact ties
    else:
        X, y, _ = ds.regression_sum(seed, n, d, 2, 0.5)
        X = ds.quantized(X, 0.25)
    return X, y


def _oracle_preds(X, y, order, method, f, k_vals, k_fold, model):
    n = X.shape[0]
    D = np.stack([O.distance_row(X, q, method, order) for q in range(n)])
    mult = 1 if method == "forward" else -1
    return O.candidate_preds(D, X, y, k_vals, k_fold, f, mult, model)


@pytest.mark.parametrize("case", TINY)
def test_oracle_preds_match_brute_force(case):
    seed, n, d, k_vals, k_fold, model, method, order, f = case
    X, y = _tiny_data(seed, n, d, model)
    got = _oracle_preds(X, y, order, method, f, k_vals, k_fold, model)
    exp = brute_preds(X, y, order, method, f, k_vals, k_fold, model)
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_planted_mistakes_are_caught():
    """A reversed tie rule or LOO instead of k-fold must change the brute force."""
    seed, n, d, k_vals, k_fold, model, method, order, f = TINY[2]
    X, y = _tiny_data(seed, n, d, model)
    ok = _oracle_preds(X, y, order, method, f, k_vals, k_fold, model)
    wrong_tie = brute_preds(X, y, order, method, f, k_vals, k_fold, model, tie_rule="reversed")
    wrong_fold = brute_preds(X, y, order, method, f, k_vals, k_fold, model, drop_fold=True)
    assert not np.array_equal(ok, wrong_tie)
    assert not np.array_equal(ok, wrong_fold)


def test_distance_matrix_closed_form():
    X, _, _ = ds.regression_sum(40, 37, 7, 2, 0.5)
    D = O.distance_matrix(X, threads=2)
    assert np.all(np.diag(D) == 0)
    np.testing.assert_array_equal(D, D.T)
    for q in (0, 5, 36):
        np.testing.assert_array_equal(D[q], brute_row(X, q, [], "backward"))
    exact = ((X[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    assert np.max(np.abs(D - exact) / np.maximum(exact, 1e-30)) < 1e-5


def test_update_is_in_order_float_sum():
    X, _, _ = ds.regression_sum(41, 20, 6, 2, 0.5)
    D = np.zeros((20, 20), np.float32)
    for f in (3, 0, 5):
        O.update(X, D, f, +1, threads=1)
    for q in range(20):
        np.testing.assert_array_equal(D[q], brute_row(X, q, [3, 0, 5], "forward"))


# ------------------------------------------------------- known answers

def _vectors():
    with open(os.path.join(GOLDEN, "losses_votes.json")) as fh:
        rec = json.load(fh)
    sets = {k: [np.frombuffer(base64.b64decode(s), np.float32) for s in v]
            for k, v in rec["sets"].items()}
    return rec, sets


def test_losses_match_reference_known_answers():
    rec, sets = _vectors()
    for item in rec["losses"]:
        y, p = sets[item["set"]]
        got = np.float32(O.loss(y, p, item["loss"])).view(np.uint32)
        assert int(got) == item["value_bits"], item


def test_votes_match_reference_known_answers():
    rec, _ = _vectors()
    for item in rec["votes"]:
        lab = np.array(item["labels"], np.float32)
        assert O.mode(lab) == item["mode"], item
        assert brute_mode(lab) == np.float32(item["mode"]), item


def test_loss_closed_forms():
    y = np.array([1, -1, 1, 1, -1, 0.5, 2.0, -3.0], np.float32)
    p = np.array([1, 1, -1, 1, -1, 0.0, 1.25, -3.0], np.float32)
    # zero_one: error iff the float difference truncates to a nonzero int
    n_err = sum(int(a - b) != 0 for a, b in zip(y, p))
    assert O.loss(y, p, "zero_one") == np.float32(n_err) / np.float32(len(y))
    s = np.float32(0)
    for a, b in zip(y, p):
        t = np.float32(a - b)
        s = np.float32(s + np.float32(t * t))
    expect = np.float32(np.sqrt(np.float64(np.float32(s / np.float32(len(y))))))
    assert O.loss(y, p, "rmse") == expect
    # std keeps only the last squared deviation (loss_functions.c:141)
    m = np.float32(0)
    for a, b in zip(y, p):
        m = np.float32(m + np.float32(a - b))
    m = np.float32(m / np.float32(len(y)))
    last = np.float32(np.float32(y[-1] - p[-1]) - m)
    expect = np.float32(np.sqrt(np.float64(np.float32(np.float32(last * last) / np.float32(7)))))
    assert O.loss(y, p, "std") == expect


@pytest.mark.skipif(not O.ref_available(), reason="reference sources not built here")
@pytest.mark.parametrize("seed", [51, 52, 53])
def test_oracle_matches_reference_random(seed):
    X, y, _ = ds.classification_argmax(seed, 90, 9, 3, 3, 0.1)
    X = ds.quantized(X, 0.5)
    for method in ("forward", "backward"):
        args = (X, y, [1, 2, 4, 5], 4, 9, ["zero_one", "rmse"], seed == 53,
                "classification", method)
        assert_same_result(O.select(*args), O.ref_select(*args), f"{seed} {method}")


# ------------------------------------------------- FP64 distance rows

def _rel_err(got, ref):
    nz = ref > 0
    return float((np.abs(got[nz] - ref[nz]) / ref[nz]).max())


def _f64_check(row, X, q, feats):
    """The float64 reference for a row: numpy's sum over the subset's squared
    float64 differences (util.c:10-17 evaluated in exact-enough arithmetic)."""
    Xs = X[:, feats].astype(np.float64)
    ref = ((Xs[q] - Xs) ** 2).sum(axis=1)
    return _rel_err(np.asarray(row, np.float64), ref) < 1e-13


@pytest.mark.parametrize("feats", [[7, 3, 33, 0, 19], list(range(40)), [39, 38, 2]],
                         ids=["forward-subset", "all-features", "backward-tail"])
def test_distance_row_f64_matches_numpy_float64(feats):
    """oracle_distance_row_f64 is the FP64 yardstick of the MFMA Gram and of
    the compensated update: it must agree with a float64 numpy computation
    to round-off, on features of scales 10^-2..10^2."""
    Z = ds.gaussian_matrix(93, 300, 40)
    scales = ds.exp2(np.log2(10.0) * (4.0 * ds.uniforms(94, 0, 40) - 2.0))
    X = (Z.astype(np.float64) * scales[None, :]).astype(np.float32)
    for q in (0, 151, 299):
        row = O.distance_row_f64(X, q, feats)
        assert row[q] == 0.0
        assert _f64_check(row, X, q, feats)


def test_distance_row_f64_planted_mistakes_fail():
    """Plausible slips in an FP64 reference -- float32 accumulation, a dropped
    feature, a transposed operand (row of another query) -- fail the check."""
    Z = ds.gaussian_matrix(95, 200, 30)
    scales = ds.exp2(np.log2(10.0) * (4.0 * ds.uniforms(96, 0, 30) - 2.0))
    X = (Z.astype(np.float64) * scales[None, :]).astype(np.float32)
    feats, q = list(range(30)), 17
    f32 = np.zeros(200, np.float32)
    for f in feats:
        t = (X[q, f] - X[:, f]).astype(np.float32)
        f32 = (f32 + t * t).astype(np.float32)
    assert not _f64_check(f32.astype(np.float64), X, q, feats)
    assert not _f64_check(O.distance_row_f64(X, q, feats[:-1]), X, q, feats)
    assert not _f64_check(O.distance_row_f64(X, q + 1, feats), X, q, feats)
