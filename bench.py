"""Benchmark of the gfx950 squared-Euclidean path.

Default workload (one GPU): brute-force KNeighborsClassifier predict.
Database 1,000,000 x 128 float32 from a mixture of 64 Gaussians (means
U[-5,5], per-component std U[0.5,2]), class = component mod 10; 100,000
queries from the same mixture; k = 10, uniform weights.  A step is one
predict over all queries, with database and queries already in HBM.  The
mixed-precision form is timed for `value`; the plain f64 form is timed
beside it; both must give the same neighbour indices and labels, and a
sample of queries is checked against the CPU oracle, which is also timed as
the CPU baseline on the host's cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload knn|kmeans]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "bf16": 2.5e15, "hbm": 8.0e12}   # MI355X spec: FLOP/s, B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def knn_data(N=1_000_000, d=128, Q=100_000, comps=64, seed=0):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-5, 5, size=(comps, d)).astype(np.float32)
    sd = rng.uniform(0.5, 2.0, size=comps).astype(np.float32)

    def draw(n):
        c = rng.integers(0, comps, size=n)
        x = rng.standard_normal((n, d), dtype=np.float32)
        x *= sd[c, None]
        x += means[c]
        return x, c

    X, cx = draw(N)
    Qx, _ = draw(Q)
    return X, (cx % 10).astype(np.int64), Qx


def cpu_baseline_knn(X, y, Qx, gpu_labels, gpu_idx, k):
    """Times the oracle (the CPU restatement of the reference) on a bounded
    sample of the queries and checks the GPU answers on that sample."""
    from oracle import oracle as O
    threads = O.max_threads()
    X64 = X.astype(np.float64)
    t0 = time.perf_counter()
    probe = 32
    O.knn_classify(X64, y, Qx[:probe].astype(np.float64), k, nthreads=threads)
    rate = probe / (time.perf_counter() - t0)
    m = int(min(4096, max(probe, rate * 15.0)))
    Qs = Qx[:m].astype(np.float64)
    t0 = time.perf_counter()
    lab = O.knn_classify(X64, y, Qs, k, nthreads=threads)
    dt = time.perf_counter() - t0
    _, idx = O.kneighbors(X64, Qs[:256], k, nthreads=threads)
    exact = bool((lab == gpu_labels[:m]).all() and (idx == gpu_idx[:256]).all())
    return {"value": m / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"first {m} of {len(Qx)} queries against the full 1,000,000-row database, "
                      f"oracle/ref_cpu.c with {threads} OpenMP threads, {dt:.1f} s; "
                      "rate extrapolated to the whole query set"}, exact


def time_predict(clf, qd, steps, warmup, sync):
    reps = []
    for _ in range(warmup):
        clf.predict(qd)
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = clf.predict(qd)
        reps.append(clf.report)
    sync()
    dt = time.perf_counter() - t0
    return out, dt, reps


def run_knn(args):
    import ferrolearn_amd as F

    k = 10
    log("generating data ...")
    X, y, Qx = knn_data(Q=args.queries)
    N, d = X.shape
    Q = Qx.shape[0]
    qd = F.to_device(Qx)          # queries resident in HBM before timing
    forms = {}
    base = None
    for form in args.forms:
        log(f"form {form}: fit + warmup")
        clf = F.KNeighborsClassifier().with_n_neighbors(k).with_precision(form).fit(X, y)
        labels, dt, reps = time_predict(clf, qd, args.steps, args.warmup, F.device.synchronize)
        dist, idx = clf.kneighbors(qd)
        ms_dist = float(np.mean([r["ms_distance"] for r in reps]))
        forms[form] = {
            "queries_per_s": Q * args.steps / dt,
            "ms_per_step": dt * 1e3 / args.steps,
            "ms_distance_kernel": ms_dist,
            "ms_finalize": float(np.mean([r["ms_finalize"] for r in reps])),
            "ms_fallback": float(np.mean([r["ms_fallback"] for r in reps])),
            "candidates_per_query": reps[-1]["n_candidates"] / Q,
            "fallback_queries": int(reps[-1]["n_fallback"]),
            "labels": labels.numpy(), "idx": idx.numpy(),
        }
        log(f"form {form}: {forms[form]['queries_per_s']:.0f} q/s, "
            f"{forms[form]['ms_per_step']:.2f} ms/step, kernel {ms_dist:.2f} ms")
        del clf
    ref = forms["fp64"] if "fp64" in forms else None
    head = args.forms[0]
    h = forms[head]
    agree = {f: bool((v["idx"] == ref["idx"]).all() and (v["labels"] == ref["labels"]).all())
             for f, v in forms.items()} if ref else {}
    flops = 2.0 * Q * N * d
    peak_key = "f64" if head in ("fp64", "exact") else "bf16"
    achieved = flops / (h["ms_distance_kernel"] * 1e-3)
    line = {
        "metric": "KMeans samples·iter/s + kNN queries/s, mixed vs FP64 GPU, exact labels",
        "value": h["queries_per_s"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if peak_key == "f64" else "bf16",
        "data": "synthetic",
        "config": {"workload": "knn_classify_bruteforce (configs[1])", "db_rows": N,
                   "features": d, "queries": Q, "k": k, "weights": "uniform",
                   "db_dtype": "float32", "form": head},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12,
                     "peak": PEAK[peak_key] / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK[peak_key], "traffic": None,
                     "kernel": "dist_bf16_kernel" if peak_key == "bf16" else "dist_topm_kernel",
                     "algorithmic_flops_per_launch": flops,
                     "kernel_ms": h["ms_distance_kernel"]},
        "forms": {f: {kk: vv for kk, vv in v.items() if kk not in ("labels", "idx")}
                  for f, v in forms.items()},
        "exact_vs_fp64_form": agree,
    }
    if ref is not None and head != "fp64":
        line["speedup_vs_fp64_form"] = h["queries_per_s"] / ref["queries_per_s"]
    if not args.no_cpu:
        log("cpu baseline ...")
        cb, ok = cpu_baseline_knn(X, y, Qx, h["labels"], h["idx"], k)
        line["cpu_baseline"] = cb
        line["exact_vs_oracle_sample"] = ok
    return line


def copy_calibration(mib=512):
    """A plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    counters give the scale that turns those counters into HBM bytes."""
    import ferrolearn_amd as F
    from ferrolearn_amd import _lib as L
    import ctypes as C
    n = mib * (1 << 20) // 4
    a = F.to_device(np.ones(n, np.float32))
    b = F.device.empty((n,), np.float32)
    for _ in range(3):
        L.lib().fl_memcpy_device(C.c_void_p(b.data_ptr()), C.c_void_p(a.data_ptr()), n * 4)
    F.device.synchronize()
    log(f"copied {mib} MiB three times")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="knn", choices=["knn", "copy"])
    ap.add_argument("--forms", default="bf16x3,fp64",
                    help="comma list; the first is reported as value (fp64, exact, bf16, bf16x3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--queries", type=int, default=100_000,
                    help="query count (100,000 is the benchmark; fewer only for profiling runs)")
    args = ap.parse_args()
    args.forms = [f for f in args.forms.split(",") if f]
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        copy_calibration()
        return
    line = run_knn(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
