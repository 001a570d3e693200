"""Benchmark: greedy kNN wrapper feature selection on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c0|c1|c3|c4]
                    [--form precision|plain] [--no-plain] [--no-cpu-baseline]
    python bench.py --copy-calibration BYTES   # a device copy, for counter scaling
    python bench.py --gram c2|c4               # the full-d distance matrix alone

One bench step = one complete selection over the workload (every selection
step: candidate gather, kNN search over all candidates with leave-one-out,
losses, arg-min, distance update), with X and y already resident in HBM.  The
default workload is forward selection on n=65536 points, d=512 float32
features with scales 10^U(-2,2), k=20, 2 classes, to 32 features, in the
precision form, followed by one plain-form selection of the same problem
(its wall time and whether the two forms' results are bit-identical).  c0:
n=2000, d=50, k=10 to 10 features.  c1: n=20000, d=200, k=10, 3 classes, to
20 features.  c3: 64 independent problems (n 1000..16000, d 32..256, k
5/10/20, to 8 features each) in one batched call per bench step, beside 64
sequential plain-form selections.  c4: backward elimination, n=32768,
d=1024, k=15, regression, to 16 features.

Printed: one JSON line with the whole-job throughput in (query x candidate)
kNN evaluations per second, the selection wall time, the dominant search
kernel's roofline (algorithmic bytes: each query's D row once plus 4 bytes
per (point, candidate) value the search evaluates, over the kernel's time
from HIP events on the engine's stream; HBM traffic from the committed
rocprofv3 counter summary when one exists) and the CPU baselines: the oracle
on a bounded sample of the same workload, and the reference's own compiled
CPU module beside the oracle on the n=2000 workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "forward-selection wall s + (query x candidate) kNN evals/s; % HBM roofline"
LIST_SIZES = [2, 4, 5, 8, 10, 15, 16, 20, 24, 32, 50, 64]


def list_size(kmax: int) -> int:
    return next(v for v in LIST_SIZES if kmax <= v)


def traffic_per_launch(workload: str):
    """(HBM bytes per search launch, source) from the committed rocprofv3
    counter summary profiles/traffic_search_<workload>.md, or (None, None).
    The summary averages over the pruned-search launches of the counted
    selection and names the commit its kernels were built from."""
    path = os.path.join(ROOT, "profiles", f"traffic_search_{workload}.md")
    if not os.path.exists(path):
        return None, None
    text = open(path).read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+-]+)", text)
    c = re.search(r"measured_on:\s*(\S+)", text)
    src = f"profiles/traffic_search_{workload}.md" + (f" (kernels of {c.group(1)})" if c else "")
    return (float(m.group(1)), src) if m else (None, None)


def cpu_baseline(X, y, kv, k_fold, model_type, method, target_s: float = 12.0) -> dict:
    """Oracle (oracle/oracle.c) rate on a bounded sample: the first selection
    step for a prefix of the queries over every candidate."""
    from oracle import oracle as O
    O.build()
    n, d = X.shape
    threads = max(1, min(16, os.cpu_count() or 1))
    mult = 1 if method == 1 else -1

    def run(nq):
        q = np.arange(nq, dtype=np.int64)
        if mult > 0:
            Drows = np.zeros((nq, n), np.float32)
        else:  # backward start: full-d distance rows of the sampled queries
            Drows = np.stack([O.distance_row(X, int(i), "backward", []) for i in q])
        t = time.perf_counter()
        for f in range(d):
            O.candidate_preds(Drows, X, y, kv, k_fold, f, mult, model_type, queries=q,
                              threads=threads)
        return time.perf_counter() - t

    nq = 16
    dt = run(nq)
    while dt < 1.0 and nq < n:  # grow until the rate estimate is stable
        nq = min(n, nq * 4)
        dt = run(nq)
    if dt < 0.5 * target_s and nq < n:
        nq = int(min(n, nq * target_s / max(dt, 1e-3)))
        dt = run(nq)
    step = "first forward step (D=0)" if mult > 0 else "first backward step (full D)"
    return {"value": nq * d / dt, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle, {step}: {nq} queries x {d} candidates x {n} points, "
                      f"{dt:.2f} s on {threads} threads"}


def roofline(alg_bytes: float, ms: float, launches: int, kernel: str, traffic, unit_note: str):
    achieved = alg_bytes / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
           "launches": launches, "avg_launch_ms": ms / max(launches, 1),
           "alg_bytes_per_launch": alg_bytes / max(launches, 1), "alg_bytes": unit_note}
    if traffic is not None and ms > 0:
        phys = traffic * launches / (ms / 1e3) / 1e9
        out["physical_GBps"] = phys
        out["physical_frac"] = phys / HBM_PEAK_GBS
    return out


def reference_vs_port_c0(timeout_s: float = 240.0):
    """The reference's own compiled CPU module (oracle/_ref, built from the
    reference sources; single-threaded by construction) and the oracle on one
    thread, each running the whole n=2000 d=50 forward selection, timed in a
    child process (the reference prints to stdout on some paths).  None where
    the reference module is not built."""
    from oracle import oracle as O
    if not O.ref_available():
        return None
    code = (
        "import json, time, sys\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "from oracle import oracle as O\n"
        "from speedynn_amd import datasets as ds\n"
        "w = ds.config('c0')\n"
        "args = (w['X'], w['y'], w['k_vals'], w['num_features'], w['k_fold'], "
        "w['loss_functions'], False, w['model_type'], w['method'])\n"
        "t = time.perf_counter(); r = O.ref_select(*args); tr = time.perf_counter() - t\n"
        "t = time.perf_counter(); p = O.select(*args, threads=1); tp = time.perf_counter() - t\n"
        "same = r['selected_ordering'].tolist() == p['selected_ordering'].tolist() and "
        "r['selected_errors'].tobytes() == p['selected_errors'].tobytes()\n"
        "print('@@' + json.dumps(dict(reference_s=tr, port_1thread_s=tp, same=same)))\n")
    try:
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                             timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return None
    for line in out.stdout.splitlines():
        if line.startswith("@@"):
            return json.loads(line[2:])
    return None


def pruned_kernel_names(pruned_kernel: int, K: int, mult: int, nblk: int) -> str:
    W = 4 if nblk <= 4 else 8 if nblk <= 8 else 16
    names = []  # the kernels the autotuner kept for this workload's steps
    if pruned_kernel & 1:
        names.append(f"k_search_pruned<{K},{mult}>")
    if pruned_kernel & 2:
        names.append(f"k_search_pruned_shared<{K},{mult},{W}>")
    if pruned_kernel & 4:  # the same kernel, two workgroups per query
        names.append(f"k_search_pruned_shared<{K},{mult},{4 if nblk <= 8 else 8}> x2")
    if pruned_kernel & 8:  # the second step's walk along a sorted column
        names.append(f"k_search_walk<{K}>")
    for bit, width in ((16, 4), (32, 8), (64, 16)):  # the tuner's workgroup widths
        if pruned_kernel & bit:
            parts = (nblk + width - 1) // width
            names.append(f"k_search_pruned_shared<{K},{mult},{width}>"
                         + (f" x{parts}" if parts > 1 else ""))
    return " + ".join(names) or f"k_search_pruned<{K},{mult}>"


def bench_single(args, hip, eng, form: int) -> dict:
    from speedynn_amd import datasets as ds
    w = ds.config(args.workload)
    X, y = w["X"], w["y"]
    n, d = X.shape
    method = 1 if w["method"] == "forward" else 0
    mt = 1 if w["model_type"] == "classification" else 0
    loss = 4 if mt == 1 else 0
    xp, yp = hip.upload(np.ascontiguousarray(X)), hip.upload(np.ascontiguousarray(y))
    kw = dict(method=method, model_type=mt, k_fold=w["k_fold"], num_features=w["num_features"],
              device_inputs=(xp, yp, n, d))

    def once(f=form, graph=False):
        return eng.select(None, None, w["k_vals"], [loss], form=f, use_graph=graph, **kw)

    cold = None  # the process's first selection: the autotuner times its kernels here
    for i in range(args.warmup):
        t_w = time.perf_counter()
        once()
        hip.sync()
        if i == 0:
            cold = time.perf_counter() - t_w
    hip.sync()
    t0 = time.perf_counter()
    results = [once() for _ in range(args.steps)]
    hip.sync()
    wall = time.perf_counter() - t0
    r = results[-1]
    steps_sel = r["steps"]
    evals = sum(x["evals"] for x in results)
    npf = n // w["k_fold"]

    # dominant kernel: the pruned search of the steps after a forward
    # selection's first (that one runs over sorted columns), else the dense
    # search.  Algorithmic bytes per launch: each query's D row once (4n) plus
    # 4 bytes per candidate value the search evaluates -- the (point,
    # candidate) pairs it visits, counted by the kernels; points that
    # exact pruning skips move nothing.
    K = list_size(max(w["k_vals"]))
    first = 1 if (form == 1 and method == 1) else 0
    launches = args.steps * (steps_sel - first)
    ms_dom = sum(x["ms_search"] - (x["ms_search_first"] if first else 0.0) for x in results)
    full = args.steps * sum((d - t) * n * (n - npf) for t in range(first, steps_sel))
    if form == 1:
        kernel = pruned_kernel_names(r["pruned_kernel"], K, 1 if method == 1 else -1,
                                     (d + 63) // 64)
        visited = full - sum(x["pruned_point_evals"] for x in results)
    else:
        kernel = f"k_search_dense<{K}>"
        visited = full
    alg = 4.0 * n * n * launches + 4.0 * visited
    traffic, traffic_src = traffic_per_launch(args.workload) if form == 1 else (None, None)
    rf = roofline(alg, ms_dom, launches, kernel, traffic,
                  "4*n per query (its D row) + 4 B per (point, candidate) value the search "
                  "evaluates; pruned points move nothing")
    if traffic_src:
        rf["traffic_source"] = traffic_src

    parity = None
    gold = os.path.join(ROOT, "tests", "golden", f"{args.workload}_oracle.json")
    if args.workload == "c0":
        gold = os.path.join(ROOT, "tests", "golden", "c0.json")
    if os.path.exists(gold):
        with open(gold) as fh:
            exp = json.load(fh)["expected"]
        parity = bool(np.array_equal(r["selected_ordering"], exp["selected_ordering"]) and
                      np.array_equal(r["selected_errors"].view(np.uint32),
                                     np.array(exp["selected_errors_bits"], np.uint32)))
    line = {
        "metric": METRIC, "value": evals / wall, "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (deterministic splitmix64 + Box-Muller generator)",
        "config": {"workload": f"{args.workload}: {w['method']} selection n={n} d={d} "
                               f"k={w['k_vals'][0]} LOO {'zero_one' if mt else 'rmse'} to "
                               f"{w['num_features']} features",
                   "n": n, "d": d, "k": w["k_vals"][0], "features": w["num_features"],
                   "form": args.form, "parallelism": "single GPU"},
        "selection_wall_s": wall / args.steps,
        "first_selection_wall_s": cold,
        "selection_steps": steps_sel,
        "search_ms_per_selection": sum(x["ms_search"] for x in results) / args.steps,
        "first_step_search_ms": sum(x["ms_search_first"] for x in results) / args.steps,
        "bound_ms_per_selection": sum(x["ms_bound"] for x in results) / args.steps,
        "update_ms_per_selection": sum(x["ms_update"] for x in results) / args.steps,
        "tally_ms_per_selection": sum(x["ms_tally"] for x in results) / args.steps,
        "distance_ms_per_selection": sum(x["ms_distance"] for x in results) / args.steps,
        # share of the (point, candidate) evaluations of the pruned steps that
        # exact pruning skipped
        "pruned_point_fraction": (1.0 - visited / full) if (form == 1 and full) else 0.0,
        "roofline": rf,
        "parity_vs_golden": parity,
        # forward: features in the order chosen; backward: the first 32 removed
        "selected": (r["selected_ordering"][:steps_sel].tolist() if method == 1 else
                     r["selected_ordering"][::-1][:min(steps_sel, 32)].tolist()),
    }
    if not args.no_graph:
        # the same selection with its steps replayed from hipGraphs (one call
        # captures them), outside the timed region; graph mode keeps no phase
        # times, so the roofline above comes from the stream-issued steps
        once(graph=True)
        hip.sync()
        ng = min(args.steps, 3)
        t1 = time.perf_counter()
        gr = [once(graph=True) for _ in range(ng)]
        hip.sync()
        line["graph_selection_wall_s"] = (time.perf_counter() - t1) / ng
        line["graph_equals_stream"] = all(
            np.array_equal(g["selected_ordering"], r["selected_ordering"]) and
            np.array_equal(g["selected_errors"].view(np.uint32),
                           r["selected_errors"].view(np.uint32)) for g in gr)
    if form == 1 and not args.no_plain:
        # the plain FP32 form (dense search every step) once, outside the
        # timed region: its wall time and whether the precision form's
        # selection and per-step errors are bit-identical to it
        hip.sync()
        t1 = time.perf_counter()
        pl = once(0)
        hip.sync()
        plain = time.perf_counter() - t1
        line["plain_selection_wall_s"] = plain
        line["speedup_vs_plain"] = plain / (wall / args.steps)
        line["precision_equals_plain"] = bool(
            np.array_equal(pl["selected_ordering"], r["selected_ordering"]) and
            np.array_equal(pl["selected_errors"].view(np.uint32),
                           r["selected_errors"].view(np.uint32)))
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(X, y, w["k_vals"], w["k_fold"], mt, method)
        rc0 = reference_vs_port_c0()
        if rc0:
            line["cpu_baseline"]["reference_c0"] = {
                "reference_s": rc0["reference_s"], "port_1thread_s": rc0["port_1thread_s"],
                "same_selection": rc0["same"],
                "note": "whole n=2000 d=50 forward selection to 10 features: the reference's "
                        "own compiled CPU module (single-threaded) beside the oracle on one "
                        "thread; the reference cannot run this workload (n > 15000 exits)"}
    hip.free(xp)
    hip.free(yp)
    return line


def bench_batch(args, hip, eng, form: int) -> dict:
    from speedynn_amd import datasets as ds
    probs, ptrs = [], []
    for p in range(64):
        w = ds.batch_problem(p)
        n, d = w["X"].shape
        xp, yp = hip.upload(np.ascontiguousarray(w["X"])), hip.upload(np.ascontiguousarray(w["y"]))
        ptrs += [xp, yp]
        probs.append(dict(device_inputs=(xp, yp, n, d), X=w["X"], y=w["y"], k_vals=w["k_vals"],
                          loss_codes=[4], method=1, model_type=1, k_fold=w["k_fold"],
                          num_features=w["num_features"]))

    def once(graph=True):
        return eng.select_batch(probs, form=form, use_graph=graph)

    for _ in range(args.warmup):
        once()
    hip.sync()
    t0 = time.perf_counter()
    runs = [once() for _ in range(args.steps)]
    hip.sync()
    wall = time.perf_counter() - t0
    got, sm = runs[-1]
    evals = sum(s["evals"] for _, s in runs)

    # phase split from one more call without graphs (not part of the timed value)
    _, split = once(graph=False)
    first = 1 if form == 1 else 0
    # algorithmic bytes of the dominant kernel's launches: each problem's D
    # rows once per step (4 n_p^2) plus 4 bytes per (point, candidate) value
    # the search evaluates (pruned points move nothing)
    rows, full = 0.0, 0
    for p, g in zip(probs, got):
        n, d = p["device_inputs"][2], p["device_inputs"][3]
        npf = n // p["k_fold"]
        rows += 4.0 * n * n * (g["steps"] - first)
        full += sum((d - t) * n * (n - npf) for t in range(first, g["steps"]))
    visited = full - split["pruned_point_evals"] if form == 1 else full
    dom_bytes = rows + 4.0 * visited
    ms_dom = split["ms_search"] - (split["ms_search_first"] if first else 0.0)
    sizes = sorted({list_size(max(p["k_vals"])) for p in probs})
    launches = len(sizes) * (max(g["steps"] for g in got) - first)  # one per list size and step
    family = "k_batch_pruned" if form == 1 else "k_batch_dense"
    kernel = f"{family}<K> (K = {', '.join(map(str, sizes))})"
    rf = roofline(dom_bytes, ms_dom, launches, kernel,
                  traffic_per_launch("c3")[0] if form == 1 else None,
                  "4*n_p per query of problem p (its D row) + 4 B per (point, candidate) "
                  "value the search evaluates; pruned points move nothing")

    # the comparison: 64 sequential plain-form selections of the same problems
    seq = None
    if not args.no_sequential:
        hip.sync()
        t1 = time.perf_counter()
        for p in probs:
            xp, yp, n, d = p["device_inputs"]
            eng.select(None, None, p["k_vals"], [4], method=1, model_type=1, k_fold=p["k_fold"],
                       num_features=p["num_features"], form=0, device_inputs=(xp, yp, n, d))
        hip.sync()
        seq = time.perf_counter() - t1

    line = {
        "metric": METRIC, "value": evals / wall, "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (deterministic splitmix64 + Box-Muller generator)",
        "config": {"workload": "c3: 64 forward selections in one batched call, n 1000..16000, "
                               "d 32..256, k 5/10/20, LOO zero_one to 8 features",
                   "problems": 64, "form": args.form, "graphs": True,
                   "parallelism": "single GPU"},
        "selection_wall_s": wall / args.steps,
        "sequential_plain_s": seq,
        "speedup_vs_sequential_plain": (seq / (wall / args.steps)) if seq else None,
        "search_ms_per_batch": split["ms_search"],
        "first_step_search_ms": split["ms_search_first"],
        "bound_ms_per_batch": split["ms_bound"],
        "roofline": rf,
    }
    if not args.no_cpu_baseline:
        p0 = probs[0]
        line["cpu_baseline"] = cpu_baseline(p0["X"], p0["y"], p0["k_vals"], p0["k_fold"], 1, 1)
        line["cpu_baseline"]["sample"] += " (problem 0 of 64)"
    for p in ptrs:
        hip.free(p)
    return line


def bench_gram(args, hip, eng) -> dict:
    """(a) alone: the full-d distance matrix of a workload's X, exact FP32 in
    the reference's order (VALU) and the BF16x3 MFMA Gram; best of 3 after a
    warm-up, HIP events around the matrix kernels."""
    from speedynn_amd import datasets as ds
    w = ds.config(args.gram)
    X = np.ascontiguousarray(w["X"])
    n, d = X.shape
    xp, dp = hip.upload(X), hip.alloc(4 * n * n)
    times = {}
    for form, name in ((2, "mfma"), (0, "exact")):
        eng.distance_matrix(xp, n, d, form, dp)
        times[name] = min(eng.distance_matrix(xp, n, d, form, dp) for _ in range(3))
    hip.free(xp)
    hip.free(dp)
    flops = 2.0 * n * n * d  # algorithmic: one multiply-add per (pair, feature)
    # products the matrix cores execute: the upper triangle of square tiles
    # (256 points from n = 8192 on, else 128; diagonal tiles whole), features
    # padded to 16, six split products each
    tile = 256 if n >= 8192 else 128
    nb = (n + tile - 1) // tile
    executed = 6 * 2.0 * (nb * (nb + 1) // 2) * tile * tile * ((d + 15) // 16 * 16)
    t_m, t_e = times["mfma"] / 1e3, times["exact"] / 1e3
    return {"gram": args.gram, "n": n, "d": d,
            "mfma_ms": times["mfma"], "exact_ms": times["exact"],
            "mfma_alg_TFLOPs": flops / t_m / 1e12,
            "mfma_executed_TFLOPs": executed / t_m / 1e12,
            "mfma_util": executed / t_m / 2.5e15,
            "mfma_write_GBps": 4.0 * n * n / t_m / 1e9,
            "exact_alg_TFLOPs": flops / t_e / 1e12,
            "note": "mfma_util: the bf16 products the matrix cores execute (six split "
                    "products per feature pair, one triangle of tiles, diagonal tiles whole, "
                    "features padded to 16) over the time, against the 2.5 PFLOP/s dense "
                    "BF16 peak; mfma_alg_TFLOPs: 2n^2d over the time; the D write is 4n^2 "
                    "bytes; times include the mean and split kernels"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2", "c3", "c4"])
    ap.add_argument("--form", default="precision", choices=["precision", "plain"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sequential", action="store_true", help="c3: skip the sequential runs")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one plain-form selection beside the precision form")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the selections replayed from step graphs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run device-to-device copies of BYTES (counter calibration) and exit")
    ap.add_argument("--gram", choices=["c1", "c2", "c4"], default=None,
                    help="time only the full-d distance matrix of a workload and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the selection "
              "engine runs on a single device.", file=sys.stderr)
        return 2

    from speedynn_amd._lib import DeviceMemory
    hip = DeviceMemory()
    if args.copy_calibration:
        a, b = hip.alloc(args.copy_calibration), hip.alloc(args.copy_calibration)
        hip.copy_d2d(b, a, args.copy_calibration)
        hip.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            hip.copy_d2d(b, a, args.copy_calibration)
        hip.sync()
        dt = (time.perf_counter() - t0) / 3
        hip.free(a)
        hip.free(b)
        # read + write bytes per copy over its wall time: what a streaming
        # read-modify-write kernel can expect from this device's HBM
        print(json.dumps({"copy_bytes": args.copy_calibration, "copies": 4,
                          "copy_read_write_GBps": 2 * args.copy_calibration / dt / 1e9}),
              flush=True)
        return 0

    from speedynn_amd import _lib
    eng = _lib.Engine(0)
    form = _lib.FORMS[args.form]
    if args.gram:
        print(json.dumps(bench_gram(args, hip, eng)), flush=True)
        eng.close()
        return 0
    if args.workload == "c3":
        line = bench_batch(args, hip, eng, form)
    else:
        line = bench_single(args, hip, eng, form)
    eng.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
