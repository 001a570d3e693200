"""Golden selections for workloads the reference cannot run (n > 15000 makes it
exit, speedynn/lib/nearest_neighbors_feature_selection.c:35,84-87,251-254):
computed by the CPU oracle (oracle/oracle.c, itself pinned against the
reference's own outputs by the fixtures of make_golden.py).

    python tests/golden/make_oracle_golden.py c1
    python tests/golden/make_oracle_golden.py c2 3000 512 32   # c2's recipe at n=3000
    python tests/golden/make_oracle_golden.py c4 1000 320 16   # c4's recipe at n=1000, d=320

Records the selected sequence, the per-step errors (the integer LOO error
counts are errors * n) and the first-loss error of every candidate at every
step (zlib-compressed base64 of the float32 bits for the reduced recipes), plus input
checksums.
"""
from __future__ import annotations

import base64
import json
import os
import sys
import time
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from speedynn_amd import datasets as ds  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main(name: str, n: int | None = None, d: int | None = None, nf: int | None = None) -> None:
    O.build()
    w = ds.config(name) if n is None else ds.reduced(name, n, d, nf)
    t0 = time.time()
    r = O.select(w["X"], w["y"], w["k_vals"], w["num_features"], w["k_fold"],
                 w["loss_functions"], False, w["model_type"], w["method"],
                 record_candidates=True)
    secs = time.time() - t0
    n, d = w["X"].shape
    steps = r["num_final"] if w["method"] == "forward" else d - r["num_final"]
    cand = np.ascontiguousarray(r["candidate_errors"][:steps])
    inputs = dict(recipe="config", config=name, k_vals=w["k_vals"],
                  num_features=w["num_features"], k_fold=w["k_fold"],
                  loss_functions=w["loss_functions"], auto_feature=False,
                  model_type=w["model_type"], method=w["method"], n=n, d=d,
                  X_checksum=ds.checksum(w["X"]), y_checksum=ds.checksum(w["y"]))
    expected = dict(num_final=int(r["num_final"]),
                    selected_dimensions=r["selected_dimensions"].tolist(),
                    selected_ordering=r["selected_ordering"].tolist(),
                    selected_errors_bits=r["selected_errors"].view(np.uint32).tolist(),
                    selected_errors=[float(v) for v in r["selected_errors"]],
                    optimal_k=int(r["optimal_k"]))
    if w["name"] == name:
        out = f"{name}_oracle.json"
        expected["candidate_errors_bits"] = cand.view(np.uint32).tolist()
    else:
        inputs["recipe"] = "reduced"
        out = f"{name}_n{n}_d{d}_oracle.json"
        raw = cand.view(np.uint32).astype("<u4").tobytes()
        expected["candidate_errors_zb64"] = base64.b64encode(zlib.compress(raw, 9)).decode()
    rec = dict(name=w["name"], source="CPU oracle (oracle/oracle.c), %d threads, %.1f s" %
               (O.max_threads(), secs), inputs=inputs, expected=expected)
    with open(os.path.join(OUT, out), "w") as fh:
        json.dump(rec, fh)
    print(w["name"], secs, r["selected_ordering"][:steps].tolist() if w["method"] == "forward"
          else r["selected_ordering"][::-1][:steps].tolist(),
          (r["selected_errors"][:steps] * n).round().astype(int).tolist()[:8])


if __name__ == "__main__":
    a = sys.argv[1:]
    main(a[0] if a else "c1", *(int(v) for v in a[1:4]))
