"""CPU baseline timing: the reference's own compiled CPU module (oracle/_ref,
built from the reference sources by `make -C oracle ref`; single-threaded by
construction) beside the oracle on 1 and 8 threads, on whole selections of
the 64-problem batched workload's smallest problems (the reference exits for
n > 15000, so it cannot run the larger workloads).  Test and measurement
infrastructure, run in a container that holds the reference sources:

    python oracle/time_reference.py [problem ...]

Prints one JSON object per problem: shape, seconds per selection for each
path, and whether the reference and the oracle chose the same features with
bit-identical per-step errors.
"""
from __future__ import annotations

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from speedynn_amd import datasets as ds  # noqa: E402

DEFAULT = [22, 13, 25, 10, 36, 23, 29, 12]  # the smallest n * n * d of the 64


def main(problems):
    O.build(with_ref=True)
    if not O.ref_available():
        raise SystemExit("the reference module could not be built")
    for p in problems:
        w = ds.batch_problem(p)
        args = (w["X"], w["y"], w["k_vals"], w["num_features"], w["k_fold"],
                w["loss_functions"], False, w["model_type"], w["method"])
        t = time.perf_counter()
        r = O.ref_select(*args)
        tr = time.perf_counter() - t
        t = time.perf_counter()
        o1 = O.select(*args, threads=1)
        t1 = time.perf_counter() - t
        t = time.perf_counter()
        O.select(*args, threads=8)
        t8 = time.perf_counter() - t
        same = (r["selected_ordering"].tolist() == o1["selected_ordering"].tolist() and
                r["selected_errors"].tobytes() == o1["selected_errors"].tobytes())
        n, d = w["X"].shape
        print(json.dumps(dict(problem=p, n=n, d=d, k=max(w["k_vals"]),
                              features=w["num_features"], reference_s=tr, oracle_1thread_s=t1,
                              oracle_8threads_s=t8, same=same)), flush=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or DEFAULT)
