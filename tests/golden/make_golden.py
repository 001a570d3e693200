"""Write golden vectors by running the reference's own compiled CPU module.

Run in a container that has the reference sources (it builds oracle/_ref from
them with ``make -C oracle ref``):

    python tests/golden/make_golden.py

Each case becomes ``tests/golden/<name>.json``: the inputs (either a generator
recipe from speedynn_amd.datasets, verified by checksum, or the raw float32
arrays in base64 for tiny cases) and the reference's getter outputs
(selected_dimensions, selected_ordering, selected_errors as float32 bit
patterns, optimal k, number of final features).  Nothing of the reference's
source is stored, only its inputs and outputs.
"""
from __future__ import annotations

import base64
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from speedynn_amd import datasets as ds  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
ALL_LOSSES = ["rmse", "rmse_norm", "mad", "mad_norm", "zero_one", "std",
              "rmse_norm_mad_norm_std", "rmse_norm_mad_norm_galaxies", "rmse_std"]


def b64(a: np.ndarray) -> str:
    return base64.b64encode(np.ascontiguousarray(a).tobytes()).decode()


def cases():
    """(name, inputs dict, X, y).  inputs carry either a recipe or raw arrays."""
    out = []

    def raw(name, X, y, **kw):
        kw.update(X_b64=b64(X.astype(np.float32)), y_b64=b64(y.astype(np.float32)),
                  n=int(X.shape[0]), d=int(X.shape[1]))
        out.append((name, kw, X.astype(np.float32), y.astype(np.float32)))

    X, y, _ = ds.classification_sign(5, 64, 8, 3, 0.1)
    raw("tiny_cls_fwd_loo", X, y, k_vals=[1, 3, 5], num_features=4, k_fold=64,
        loss_functions=["zero_one", "rmse"], auto_feature=False, model_type="classification",
        method="forward")
    raw("tiny_cls_bwd_kfold", X, y, k_vals=[2, 4], num_features=3, k_fold=10,
        loss_functions=["zero_one"], auto_feature=False, model_type="classification",
        method="backward")
    X, y, _ = ds.regression_sum(6, 100, 10, 3, 0.5)
    raw("reg_fwd_all_losses", X, y, k_vals=list(range(2, 51)), num_features=3, k_fold=10,
        loss_functions=ALL_LOSSES, auto_feature=False, model_type="regression", method="forward")
    raw("reg_bwd_loo", X, y, k_vals=[3, 7], num_features=2, k_fold=100,
        loss_functions=["rmse", "mad"], auto_feature=False, model_type="regression",
        method="backward")
    raw("reg_fwd_auto_remainder_fold", X, y, k_vals=[3, 7], num_features=2, k_fold=7,
        loss_functions=["rmse"], auto_feature=True, model_type="regression", method="forward")
    raw("reg_bwd_auto_remainder_fold", X, y, k_vals=[3, 7], num_features=2, k_fold=7,
        loss_functions=["rmse"], auto_feature=True, model_type="regression", method="backward")
    raw("reg_fwd_unsorted_kvals", X, y, k_vals=[9, 2, 5, 5, 1], num_features=4, k_fold=100,
        loss_functions=["mad", "rmse"], auto_feature=False, model_type="regression",
        method="forward")
    Xq = ds.quantized(X, 0.5)
    raw("ties_cls_fwd_loo", Xq, np.round(y), k_vals=[4, 6], num_features=4, k_fold=100,
        loss_functions=["zero_one"], auto_feature=False, model_type="classification",
        method="forward")
    raw("ties_cls_bwd_loo", Xq, np.round(y), k_vals=[2, 4, 6], num_features=3, k_fold=100,
        loss_functions=["zero_one", "rmse"], auto_feature=False, model_type="classification",
        method="backward")
    X3, y3, _ = ds.classification_argmax(7, 300, 20, 6, 3, 0.05)
    raw("cls3_fwd_even_k", X3, y3, k_vals=[1, 2, 3, 4, 10], num_features=6, k_fold=300,
        loss_functions=["zero_one", "rmse"], auto_feature=False, model_type="classification",
        method="forward")
    Xs, ys, _ = ds.regression_sum(8, 40, 6, 2, 0.5)
    raw("reg_short_lists_kfold2", Xs, ys, k_vals=[5, 25], num_features=2, k_fold=2,
        loss_functions=["rmse"], auto_feature=False, model_type="regression", method="forward")
    raw("cls_short_lists_kfold2", Xs, np.sign(ys), k_vals=[24, 30], num_features=2, k_fold=2,
        loss_functions=["zero_one"], auto_feature=False, model_type="classification",
        method="forward")
    Xt, yt, _ = ds.classification_sign(9, 3, 3, 2, 0.0)
    raw("minimal_n3", Xt, yt, k_vals=[1, 2], num_features=2, k_fold=3,
        loss_functions=["zero_one"], auto_feature=False, model_type="classification",
        method="forward")
    X5, y5, _ = ds.classification_sign(11, 777, 33, 5, 0.05)
    raw("cls_n777_fwd_auto", X5, y5, k_vals=[5, 10, 20], num_features=3, k_fold=777,
        loss_functions=["zero_one"], auto_feature=True, model_type="classification",
        method="forward")
    # lists longer than the 64 slots of the register kernels, one of them longer than
    # the eligible points (the unfilled slots keep (MAX_FLOAT, index 0), so label 0 votes)
    X7, y7, _ = ds.classification_argmax(13, 300, 20, 6, 3, 0.05)
    raw("cls3_fwd_long_lists", X7, y7, k_vals=[70, 100, 3], num_features=3, k_fold=300,
        loss_functions=["zero_one", "rmse"], auto_feature=False, model_type="classification",
        method="forward")
    X8, y8, _ = ds.regression_sum(14, 200, 12, 3, 0.5)
    raw("reg_bwd_long_lists_kfold", X8, y8, k_vals=[65, 150], num_features=9, k_fold=7,
        loss_functions=["rmse", "mad"], auto_feature=False, model_type="regression",
        method="backward")
    X9, y9, _ = ds.classification_sign(15, 90, 6, 2, 0.1)
    raw("cls_k_above_eligible", X9, y9, k_vals=[100, 89], num_features=2, k_fold=90,
        loss_functions=["zero_one"], auto_feature=False, model_type="classification",
        method="forward")
    X6, y6, _ = ds.regression_sum(12, 500, 24, 4, 0.5)
    raw("reg_n500_bwd_k15", X6, y6, k_vals=[15], num_features=4, k_fold=500,
        loss_functions=["rmse"], auto_feature=False, model_type="regression", method="backward")

    w = ds.config("c0")
    rec = dict(recipe="config", config="c0", k_vals=w["k_vals"], num_features=w["num_features"],
               k_fold=w["k_fold"], loss_functions=w["loss_functions"], auto_feature=False,
               model_type=w["model_type"], method=w["method"], n=2000, d=50,
               X_checksum=ds.checksum(w["X"]), y_checksum=ds.checksum(w["y"]))
    out.append(("c0", rec, w["X"], w["y"]))

    shapes = ds.batch_problem_shapes()
    p = int(np.argmin([n * n * d for n, d, _ in shapes]))
    w = ds.batch_problem(p)
    rec = dict(recipe="batch_problem", problem=p, k_vals=w["k_vals"],
               num_features=w["num_features"], k_fold=w["k_fold"],
               loss_functions=w["loss_functions"], auto_feature=False,
               model_type=w["model_type"], method=w["method"], n=int(w["X"].shape[0]),
               d=int(w["X"].shape[1]), X_checksum=ds.checksum(w["X"]),
               y_checksum=ds.checksum(w["y"]))
    out.append((f"c3_problem{p}", rec, w["X"], w["y"]))
    return out


def main():
    O.build(with_ref=True)
    if not O.ref_available():
        raise SystemExit("the reference module could not be built")
    only = set(sys.argv[1:])  # case names to write (default: every case)
    for name, inp, X, y in cases():
        if only and name not in only:
            continue
        args = (X, y, inp["k_vals"], inp["num_features"], inp["k_fold"], inp["loss_functions"],
                inp["auto_feature"], inp["model_type"], inp["method"])
        r = O.ref_select(*args)
        rec = dict(name=name, source="reference speedynn/lib CPU module (USE_GPU=0)",
                   inputs=inp,
                   expected=dict(num_final=int(r["num_final"]),
                                 selected_dimensions=r["selected_dimensions"].tolist(),
                                 selected_ordering=r["selected_ordering"].tolist(),
                                 selected_errors_bits=r["selected_errors"].view(np.uint32).tolist(),
                                 selected_errors=[float(v) for v in r["selected_errors"]],
                                 optimal_k=int(r["optimal_k"])))
        with open(os.path.join(OUT, f"{name}.json"), "w") as fh:
            json.dump(rec, fh, indent=1)
        print(name, r["num_final"], r["selected_ordering"][:8].tolist(), flush=True)

    # loss and vote known answers from the reference's get_loss_all / get_mode_array
    rng_y = ds.normals(21, 0, 50).astype(np.float32)
    rng_p = (rng_y + 0.3 * ds.normals(21, 1, 50)).astype(np.float32)
    cls_y = np.sign(ds.normals(22, 0, 50)).astype(np.float32)
    cls_p = np.sign(ds.normals(22, 1, 50)).astype(np.float32)
    half = np.array([0.0, 0.5, 0.99, 1.0, -0.5, -1.0, 2.0, -2.5], np.float32)
    zero = np.zeros(8, np.float32)
    losses = []
    for tag, a, b in [("reg", rng_y, rng_p), ("cls", cls_y, cls_p), ("frac", half, zero)]:
        for ln in ALL_LOSSES:
            losses.append(dict(set=tag, loss=ln, value_bits=int(np.float32(O.ref_loss(a, b, ln)).view(np.uint32))))
    votes = []
    for seq in [[1, 1, -1], [1, -1], [-1, 1], [2, 0, 0, 2], [0, 1, 2], [2, 1, 1, 0, 0],
                [0, 0, 1, 1, 2, 2, 2], [1.5, 2.5, 1.5, 2.5], [3], [0, 1, 0, 1, 1, 0]]:
        votes.append(dict(labels=seq, mode=O.ref_mode(np.array(seq, np.float32))))
    with open(os.path.join(OUT, "losses_votes.json"), "w") as fh:
        json.dump(dict(source="reference get_loss_all / get_mode_array",
                       sets=dict(reg=[b64(rng_y), b64(rng_p)], cls=[b64(cls_y), b64(cls_p)],
                                 frac=[b64(half), b64(zero)]),
                       losses=losses, votes=votes), fh, indent=1)
    print("losses/votes written")


if __name__ == "__main__":
    main()
