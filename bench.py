"""Benchmark of the gfx950 squared-Euclidean path.

Headline workload (one GPU): brute-force KNeighborsClassifier predict.
Database 1,000,000 x 128 float32 from a mixture of 64 Gaussians (means
U[-5,5], per-component std U[0.5,2]), class = component mod 10; 100,000
queries from the same mixture; k = 10, uniform weights.  A step is one
predict over all queries, with database and queries already in HBM.  The
mixed-precision form is timed for `value`; the plain f64 form is timed
beside it; both must give the same neighbour indices and labels, and a
sample of queries is checked against the CPU oracle, which is also timed as
the CPU baseline on the host's cores.

Second workload, reported in the same line under "kmeans": KMeans Lloyd on
10,000,000 x 64 float32 samples from 1,024 Gaussian blobs (means U[-20,20],
std U[0.5,3]), k = 1,024, k-means++ initial centres (the library's, seed 0),
50 iterations (tol 0), samples resident in HBM; samples*iterations/s of the
Lloyd phase for the FP8, BF16x3 and plain f64 forms, which must agree bit for
bit.

Third, under "kmeans_config1": the reference's own CPU-runnable KMeans case
(100,000 x 32 f64 blobs, k = 16, k-means++ from the reference's StdRng seed 0,
n_init 1, max_iter 100, tol 1e-4), whole fits on the GPU against the oracle's
whole fit on the host's cores, bit for bit.

Fourth and fifth, under "router" and "batched_kmeans": the learned-router
exact kNN (20M x 96 rows, 4,096 cells, 1M queries) against the brute force on
the same handle, and 256 uneven KMeans problems with 8 restarts each in one
batched call against per-problem FP64 fits.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
                  [--workload all|knn|kmeans|config1|router|batched|copy]
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

PEAK = {"f64": 78.6e12, "bf16": 2.5e15, "fp8": 5.0e15, "hbm": 8.0e12}   # MI355X spec: FLOP/s, B/s

# HBM bytes per launch of the dominant kernel at the benchmark sizes, from two
# rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch),
# each scaled by what the counter reads for a plain 512 MiB device copy
# (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0), and the MFMA busy fraction
# (SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x 1024 SIMDs)) of the same
# kernels.  The passes and their commands are summarised in
# profiles/knn_bruteforce_1Mx128_q100k_k10.md and
# profiles/kmeans_lloyd_10Mx64_k1024.md.
TRAFFIC = {
    ("knn", "bf16x3"): (5.189e6 * 2.0 + 6.64e4) * 1024,
    ("knn", "fp64"): (8.508e6 * 2.0 + 6.15e4) * 1024,
    ("kmeans", "fp8"): (1.938e6 * 2.0 + 1.29e5) * 1024,
    ("kmeans", "bf16x3"): (3.986e6 * 2.0 + 2.63e5) * 1024,
    ("kmeans", "fp64"): (1.291e6 * 2.0 + 6.25e5) * 1024,
}
MFMA_BUSY = {("knn", "bf16x3"): 0.541, ("knn", "fp64"): 0.705, ("kmeans", "fp8"): 0.240,
             ("kmeans", "bf16x3"): 0.468, ("kmeans", "fp64"): 0.573}
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of bench.py at these sizes "
                  "(profiles/), copy-calibrated x2.0 / x1.0; not measured in this run")


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
        "config": {"workload": "knn_classify_bruteforce_1Mx128_q100k_k10", "db_rows": N,
                   "features": d, "queries": Q, "k": k, "weights": "uniform",
                   "db_dtype": "float32", "form": head},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12,
                     "peak": PEAK[peak_key] / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK[peak_key],
                     "traffic": TRAFFIC.get(("knn", head)) if Q == 100_000 else None,
                     "traffic_unit": "bytes per launch (HBM, rocprofv3 counters)",
                     "traffic_source": TRAFFIC_SOURCE,
                     "mfma_busy_profiled": MFMA_BUSY.get(("knn", head)),
                     "algorithmic_bytes_per_launch": N * d * 4 + Q * (d * 4 + k * 12),
                     "kernel": "dist_bf16_kernel<16,4,3>" if peak_key == "bf16" else "dist_f64_kernel<16,32,double,float>",
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


def kmeans_data(n=10_000_000, d=64, comps=1024, seed=0):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-20, 20, size=(comps, d)).astype(np.float32)
    sd = rng.uniform(0.5, 3.0, size=comps).astype(np.float32)
    c = rng.integers(0, comps, size=n)
    X = rng.standard_normal((n, d), dtype=np.float32)
    X *= sd[c, None]
    X += means[c]
    return X


def cpu_baseline_kmeans(X, init, gpu_labels_first):
    """One Lloyd iteration of the oracle on a bounded sample of rows against
    the same initial centres (about 10 s of CPU work); also checks the
    sample's first-iteration labels against the GPU's."""
    from oracle import oracle as O
    threads = O.max_threads()
    m = 2000
    t0 = time.perf_counter()
    O.assign(X[:m].astype(np.float64), init, nthreads=threads)
    rate = m / (time.perf_counter() - t0)
    m = int(min(len(X), max(m, rate * 10.0)))
    Xs = X[:m].astype(np.float64)
    t0 = time.perf_counter()
    lab, _ = O.assign(Xs, init, nthreads=threads)
    dt = time.perf_counter() - t0
    ok = bool((lab[:len(gpu_labels_first)] == gpu_labels_first[:m]).all())
    return {"value": m / dt, "unit": "samples*iter/s", "cores": threads, "kind": "port",
            "sample": f"assignment step (the Lloyd iteration's dominant part) of the first {m} "
                      f"of {len(X)} samples against the 1,024 initial centres, oracle/ref_cpu.c "
                      f"with {threads} OpenMP threads, {dt:.1f} s"}, ok


def run_kmeans(args):
    import ferrolearn_amd as F
    k, iters = 1024, args.km_iters
    log("kmeans: generating data ...")
    X = kmeans_data(n=args.km_rows)
    n, d = X.shape
    t0 = time.perf_counter()
    init = F.kmeans_plusplus(X, k, 0)
    init_s = time.perf_counter() - t0
    log(f"kmeans: k-means++ {init_s:.2f} s")
    xd = F.to_device(X)
    forms = {}
    for form in args.km_forms:
        km = F.KMeans(k).with_n_init(1).with_tol(0.0).with_precision(form)
        km.with_max_iter(2).fit(xd, init_centers=init)          # warm-up
        F.device.synchronize()
        t0 = time.perf_counter()
        f = km.with_max_iter(iters).fit(xd, init_centers=init)
        F.device.synchronize()
        dt = time.perf_counter() - t0
        r = f.report
        forms[form] = {
            "samples_iter_per_s": n * f.n_iter() / dt, "s": dt, "n_iter": f.n_iter(),
            "ms_per_iter": dt * 1e3 / max(1, f.n_iter()),
            "ms_distance_kernel": r["ms_distance"] / max(1, f.n_iter()),
            "ms_second_filter_per_iter": r["ms_finalize"] / max(1, f.n_iter()),
            "ms_update_per_iter": r["ms_update"] / max(1, f.n_iter()),
            "rows_to_second_filter": int(r["n_second"]),
            "fallback_rows": int(r["n_fallback"]),
            "labels": f.labels().numpy(), "centers": f.cluster_centers().numpy(),
            "inertia": f.inertia(),
        }
        log(f"kmeans {form}: {forms[form]['samples_iter_per_s']:.3g} samples*iter/s, "
            f"{forms[form]['ms_per_iter']:.2f} ms/iter, distance kernel "
            f"{forms[form]['ms_distance_kernel']:.2f} ms, second filter rows "
            f"{forms[form]['rows_to_second_filter']}, exact-scan rows {forms[form]['fallback_rows']}")
    head = args.km_forms[0]
    h = forms[head]
    out = {
        "workload": f"KMeans Lloyd, {n} x {d} float32, 1024 blobs, k={k}, {iters} iterations",
        "metric": "samples*iter/s", "value": h["samples_iter_per_s"], "form": head,
        "init_kmeanspp_s": init_s,
        "forms": {f: {kk: vv for kk, vv in v.items() if kk not in ("labels", "centers")}
                  for f, v in forms.items()},
    }
    if "fp64" in forms:
        ref = forms["fp64"]
        for f, v in forms.items():
            if f == "fp64":
                continue
            o = out["forms"][f]
            o["speedup_vs_fp64_form"] = v["samples_iter_per_s"] / ref["samples_iter_per_s"]
            o["label_agreement_vs_fp64"] = float((v["labels"] == ref["labels"]).mean())
            o["inertia_rel_error_vs_fp64"] = abs(v["inertia"] - ref["inertia"]) / ref["inertia"]
            o["exact_vs_fp64_form"] = bool(
                (v["labels"] == ref["labels"]).all() and v["n_iter"] == ref["n_iter"]
                and (v["centers"].view(np.int64) == ref["centers"].view(np.int64)).all()
                and v["inertia"] == ref["inertia"])
        if head != "fp64":
            out["speedup_vs_fp64_form"] = out["forms"][head]["speedup_vs_fp64_form"]
            out["exact_vs_fp64_form"] = out["forms"][head]["exact_vs_fp64_form"]
    flops = 2.0 * n * k * d
    peak_key = {"fp64": "f64", "exact": "f64", "fp8": "fp8"}.get(head, "bf16")
    kname = {"f64": "dist_f64_kernel<0,...> (argmin form)",
             "fp8": "dist_fp8_res_kernel (FP8 E4M3 argmin, centres resident in LDS, decision fused)",
             "bf16": "dist_bf16_res_kernel<2> (BF16x3 argmin, centres resident in LDS in "
                     "chunks of 512, decision fused)"}[peak_key]
    out["roofline"] = {"bound": "mfma", "achieved": flops / (h["ms_distance_kernel"] * 1e-3) / 1e12,
                       "peak": PEAK[peak_key] / 1e12, "unit": "TFLOP/s",
                       "frac": flops / (h["ms_distance_kernel"] * 1e-3) / PEAK[peak_key],
                       "kernel": kname,
                       "kernel_ms": h["ms_distance_kernel"],
                       "algorithmic_flops_per_launch": flops,
                       "traffic": TRAFFIC.get(("kmeans", head)) if n == 10_000_000 else None,
                       "traffic_unit": "bytes per launch (HBM, rocprofv3 counters)",
                       "traffic_source": TRAFFIC_SOURCE,
                       "mfma_busy_profiled": MFMA_BUSY.get(("kmeans", head)),
                       "algorithmic_bytes_per_launch": n * (d * 4 + 4) + k * d * 8}
    if not args.no_cpu:
        log("kmeans: cpu baseline ...")
        # labels of the first iteration: one GPU iteration from the same centres
        f1 = F.KMeans(k).with_n_init(1).with_tol(0.0).with_max_iter(1).with_precision(head) \
            .fit(xd, init_centers=init)
        cb, ok = cpu_baseline_kmeans(X, init, f1.labels().numpy())
        out["cpu_baseline"] = cb
        out["exact_vs_oracle_sample"] = ok
    return out


def config1_data(n=100_000, d=32, comps=16, seed=0):
    """make_blobs-style f64 data: `comps` isotropic unit-std Gaussians with
    centres uniform in [-10, 10]^d."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10.0, 10.0, size=(comps, d))
    lab = rng.integers(0, comps, size=n)
    return c[lab] + rng.normal(size=(n, d))


def run_config1(args):
    """The reference's CPU-runnable KMeans case, whole seeded fits: 100,000 x
    32 f64 blobs, k = 16, k-means++ from the reference's StdRng (seed 0),
    n_init 1, max_iter 100, tol 1e-4.  The GPU fit (k-means++ included) is
    timed over `steps` fits; the oracle's whole fit runs once on all host
    threads; both must agree bit for bit (labels, n_iter, centres)."""
    import ferrolearn_amd as F
    from oracle import oracle as O
    X = config1_data()
    n, d = X.shape
    k, seed = 16, 0
    xd = F.to_device(X)
    km = (F.KMeans(k).with_n_init(1).with_max_iter(100).with_tol(1e-4)
          .with_rng("stdrng").with_random_state(seed))
    for _ in range(max(1, args.warmup)):
        km.fit(xd)
    F.device.synchronize()
    steps = max(5, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        f = km.fit(xd)
    F.device.synchronize()
    dt = (time.perf_counter() - t0) / steps
    labels, centers = f.labels().numpy(), f.cluster_centers().numpy()
    out = {"workload": f"KMeans Lloyd f64, {n} x {d} blobs (16 centres), k={k}, "
                       "k-means++ seeded (StdRng seed 0), n_init 1, max_iter 100, tol 1e-4",
           "metric": "samples*iter/s", "value": n * f.n_iter() / dt,
           "fits_per_s": 1.0 / dt, "ms_per_fit": dt * 1e3, "n_iter": f.n_iter(),
           "inertia": f.inertia(), "route": "direct exact folds (k <= 32)",
           "ms_lloyd_per_iter": f.report["ms_distance"] / max(1, f.n_iter())}
    threads = O.max_threads()
    t0 = time.perf_counter()
    c_o, l_o, i_o, it_o = O.kmeans_fit_stdrng(X, k, 100, 1e-4, 1, seed, nthreads=threads)
    cdt = time.perf_counter() - t0
    reps = int(min(200, max(1, 3.0 / max(cdt, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(reps):
        O.kmeans_fit_stdrng(X, k, 100, 1e-4, 1, seed, nthreads=threads)
    cdt = (time.perf_counter() - t0) / reps
    out["cpu_baseline"] = {"value": n * it_o / cdt, "unit": "samples*iter/s", "cores": threads,
                           "kind": "port", "s_per_fit": cdt,
                           "sample": f"the whole fit (k-means++ and {it_o} Lloyd iterations) "
                                     f"repeated {reps} times, oracle/ref_cpu.c with {threads} "
                                     "OpenMP threads"}
    out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    out["exact_vs_oracle"] = bool(
        it_o == f.n_iter() and (l_o == labels).all()
        and (c_o.view(np.int64) == centers.view(np.int64)).all()
        and abs(i_o - f.inertia()) <= 1e-12 * i_o)
    log(f"config1: {out['ms_per_fit']:.2f} ms per fit ({f.n_iter()} iterations), "
        f"oracle {cdt:.2f} s, exact {out['exact_vs_oracle']}")
    return out


def router_data(N, d=96, comps=4096, n_train=2_000_000, n_test=1_000_000, seed=0):
    """Mixture of `comps` anisotropic Gaussians in d dimensions: random
    rotations, axis std log-uniform in [0.2, 3], means uniform in [-10, 10].
    Returns the database (rows grouped by component, sizes multinomial),
    training queries and test queries, float32."""
    g = np.random.default_rng(seed)
    means = g.uniform(-10, 10, size=(comps, d)).astype(np.float32)
    sd = np.exp(g.uniform(np.log(0.2), np.log(3.0), size=(comps, d))).astype(np.float32)
    rot = np.linalg.qr(g.standard_normal((comps, d, d)))[0].astype(np.float32)

    def draw(n, rng):
        counts = rng.multinomial(n, np.full(comps, 1.0 / comps))
        out = np.empty((n, d), np.float32)
        o = 0
        for c in range(comps):
            m = int(counts[c])
            if m:
                z = rng.standard_normal((m, d), dtype=np.float32) * sd[c]
                out[o:o + m] = z @ rot[c].T + means[c]
                o += m
        return out

    X = draw(N, np.random.default_rng(seed + 1))
    Qt = draw(n_train, np.random.default_rng(seed + 2))
    Qe = draw(n_test, np.random.default_rng(seed + 3))
    # queries in random order (the database order does not matter)
    Qt = Qt[np.random.default_rng(seed + 4).permutation(n_train)]
    Qe = Qe[np.random.default_rng(seed + 5).permutation(n_test)]
    return X, Qt, Qe


def run_router(args):
    """Learned-router exact kNN: routed search vs the brute force on the same
    handle, with the fallback rate; results must equal the brute force's."""
    import ferrolearn_amd as F
    sc = args.router_scale
    N, comps = int(20_000_000 * sc), max(16, int(4096 * sc))
    n_train, n_test = int(2_000_000 * sc), int(1_000_000 * sc)
    k, n_probe = 10, args.n_probe
    log(f"router: generating {N} x 96 rows, {comps} components ...")
    t0 = time.perf_counter()
    X, Qt, Qe = router_data(N, comps=comps, n_train=n_train, n_test=n_test)
    log(f"router: data {time.perf_counter() - t0:.1f} s")
    y = np.arange(N) % 10
    fitted = F.KNeighborsClassifier().with_n_neighbors(k).with_precision("bf16x3").fit(X, y)
    t0 = time.perf_counter()
    router = F.CellRouter(fitted, n_cells=comps, cell_iter=10, hidden=512, seed=0)
    build_s = time.perf_counter() - t0
    log(f"router: cells built in {build_s:.1f} s")
    qt = F.to_device(Qt)
    trep = router.train(qt, n_neighbors=k, epochs=args.router_epochs, batch_size=4096,
                        learning_rate=1e-3, seed=0)
    log(f"router: trained, loss {trep['loss_first']:.3f} -> {trep['loss_last']:.3f}, "
        f"targets {trep['ms_targets'] / 1e3:.1f} s, epochs {trep['ms_train'] / 1e3:.1f} s")
    del qt
    qe = F.to_device(Qe)
    router.kneighbors(F.to_device(Qe[:4096]), n_probe=n_probe)    # warm-up
    F.device.synchronize()
    t0 = time.perf_counter()
    rd, ri = router.kneighbors(qe, n_probe=n_probe, route="network")
    F.device.synchronize()
    routed_s = time.perf_counter() - t0
    rrep = dict(router.report)
    log(f"router: routed {n_test / routed_s:.0f} q/s, rescans {rrep['n_rescan']}, "
        f"brute-force fallback {rrep['n_fallback']}")
    router.kneighbors(qe, n_probe=n_probe, route="bounds")
    brep = dict(router.report)
    # brute force on a sample of the same queries, both precisions
    ns = min(n_test, int(args.router_brute_sample))
    qs = F.to_device(Qe[:ns])
    fitted.kneighbors(F.to_device(Qe[:256]))
    F.device.synchronize()
    t0 = time.perf_counter()
    bd, bi = fitted.kneighbors(qs)
    F.device.synchronize()
    brute_s = time.perf_counter() - t0
    exact = bool((bi.numpy() == ri.numpy()[:ns]).all()
                 and (bd.numpy().view(np.int64) == rd.numpy()[:ns].view(np.int64)).all())
    n64 = min(ns, 10_000)
    fitted.set_precision("fp64")
    t0 = time.perf_counter()
    fd, fi = fitted.kneighbors(F.to_device(Qe[:n64]))
    F.device.synchronize()
    fp64_s = time.perf_counter() - t0
    exact = exact and bool((fi.numpy() == ri.numpy()[:n64]).all())
    log(f"router: brute bf16x3 {ns / brute_s:.0f} q/s, fp64 {n64 / fp64_s:.0f} q/s, exact {exact}")
    out = {
        "workload": f"learned-router exact kNN, {N} x 96 float32, {comps} anisotropic "
                    f"Gaussians, {comps} cells, router 96-512-512-{comps} bf16, "
                    f"{n_train} training / {n_test} test queries, k={k}, n_probe={n_probe}",
        "metric": "queries/s",
        "value": n_test / routed_s,
        "fallback_rate": rrep["n_fallback"] / n_test,
        "rescan_rate": rrep["n_rescan"] / n_test,
        "rescan_rate_bounds_route": brep["n_rescan"] / n_test,
        "rows_scanned_per_query": rrep["rows_scanned"] / n_test,
        "brute_bf16x3_queries_per_s": ns / brute_s,
        "brute_fp64_queries_per_s": n64 / fp64_s,
        "speedup_vs_brute_bf16x3": (n_test / routed_s) / (ns / brute_s),
        "speedup_vs_fp64_form": (n_test / routed_s) / (n64 / fp64_s),
        "exact_vs_brute_sample": exact,
        "brute_sample_queries": ns,
        "report": {kk: rrep[kk] for kk in ("ms_route", "ms_bounds", "ms_scan", "ms_merge",
                                           "ms_fallback", "ms_total")},
        "build_s": build_s,
        "train": {kk: trep[kk] for kk in ("ms_targets", "ms_train", "loss_first", "loss_last")},
        "epochs": args.router_epochs,
    }
    if not args.no_cpu:
        from oracle import oracle as O
        nc = 16
        t0 = time.perf_counter()
        od, oi = O.kneighbors(X, Qe[:nc], k)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": nc / cdt, "unit": "queries/s", "cores": O.max_threads(),
                               "kind": "port",
                               "sample": f"{nc} queries, oracle brute force, all host threads"}
        out["exact_vs_oracle_sample"] = bool((oi == ri.numpy()[:nc]).all())
    return out


def batched_data(n_problems=256, seed=0):
    """Uneven KMeans problems: n log-uniform in [1e3, 2e5], d from {8, 16,
    32, 64}, k log-uniform in [4, 256], f32 Gaussian blobs with k true
    centres (uniform in [-10, 10], unit std)."""
    g = np.random.default_rng(seed)
    out = []
    for _ in range(n_problems):
        n = int(np.exp(g.uniform(np.log(1e3), np.log(2e5))))
        d = int(g.choice([8, 16, 32, 64]))
        k = int(np.exp(g.uniform(np.log(4), np.log(256))))
        k = min(k, n)
        cen = g.uniform(-10, 10, size=(k, d)).astype(np.float32)
        X = cen[g.integers(0, k, n)] + g.standard_normal((n, d), dtype=np.float32)
        out.append((X, k))
    return out


def run_batched(args):
    """Many KMeans problems (n_init = 8 each, max_iter 100, tol 1e-4): the
    batched call (every restart of every problem on the device at once,
    graph-captured Lloyd iterations) against the plain FP64 form, one
    KMeans.fit per problem (precision fp64) on 16 concurrent host threads.
    Both must agree problem by problem, bit for bit."""
    import ferrolearn_amd as F
    from concurrent.futures import ThreadPoolExecutor
    probs = batched_data(args.batched_problems)
    xs = [F.to_device(X) for X, _ in probs]
    F.device.synchronize()

    def model(i, k, form):
        return (F.KMeans(k).with_n_init(8).with_max_iter(100).with_tol(1e-4)
                .with_random_state(i).with_precision(form))

    items = [(x, model(i, k, "fp64"), None) for i, (x, (_, k)) in enumerate(zip(xs, probs))]
    F.kmeans_fit_batched(items[:8])                     # warm-up (kernels loaded)
    F.device.synchronize()
    t0 = time.perf_counter()
    fitted, rep = F.kmeans_fit_batched(items)
    F.device.synchronize()
    bs = time.perf_counter() - t0
    log(f"batched: {bs:.3f} s, {rep['n_rows'] / bs:.3g} samples*iter/s, "
        f"{rep['n_dist_launches']} graph replays, k-means++ {rep['ms_update']:.1f} ms, "
        f"Lloyd {rep['ms_distance']:.1f} ms")
    if args.batched_no_plain:
        return {"value": rep["n_rows"] / bs, "s": bs, "ms_kmeanspp": rep["ms_update"],
                "ms_lloyd": rep["ms_distance"], "graph_replays": rep["n_dist_launches"]}
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=args.batched_workers) as ex:
        plain = list(ex.map(lambda it: it[1].fit(it[0]), items))
    F.device.synchronize()
    ps = time.perf_counter() - t0
    plain_rows = sum(int(f.report["n_rows"]) for f in plain)
    log(f"plain fp64 fits: {ps:.2f} s, {plain_rows / ps:.3g} samples*iter/s")
    exact = all(x.n_iter() == y.n_iter() and x.inertia() == y.inertia()
                and (x.labels().numpy() == y.labels().numpy()).all()
                and (x.cluster_centers().numpy().view(np.int64)
                     == y.cluster_centers().numpy().view(np.int64)).all()
                for x, y in zip(fitted, plain))
    return {
        "workload": f"batched KMeans, {len(probs)} problems (n log-U[1e3,2e5], d in "
                    "{8,16,32,64}, k log-U[4,256]), n_init=8, max_iter=100, tol=1e-4",
        "metric": "samples*iter/s",
        "note": "samples*iter counts every restart's assignments; k-means++ for all "
                "restarts is inside both timed calls",
        "value": rep["n_rows"] / bs,
        "form": "batched (exact assignment, graph-captured iterations)",
        "s": {"batched": bs, "plain_fp64_fits": ps},
        "plain_fp64_fits": f"KMeans.fit per problem, precision fp64, "
                           f"{args.batched_workers} host threads",
        "speedup_vs_fp64_form": ps / bs,
        "ms_kmeanspp": rep["ms_update"], "ms_lloyd": rep["ms_distance"],
        "graph_replays": rep["n_dist_launches"],
        "samples_iter": rep["n_rows"], "samples_iter_plain": plain_rows,
        "exact_vs_fp64_form": bool(exact),
        "total_rows": int(sum(X.shape[0] for X, _ in probs)),
    }


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
    ap.add_argument("--workload", default="all",
                    choices=["all", "knn", "kmeans", "config1", "router", "batched", "copy"])
    ap.add_argument("--batched-problems", type=int, default=256)
    ap.add_argument("--batched-workers", type=int, default=16)
    ap.add_argument("--batched-no-plain", action="store_true",
                    help="time only the batched call (profiling runs)")
    ap.add_argument("--router-scale", type=float, default=1.0,
                    help="fraction of the router workload's rows, cells and queries")
    ap.add_argument("--router-epochs", type=int, default=3)
    ap.add_argument("--router-brute-sample", type=int, default=100_000)
    ap.add_argument("--n-probe", type=int, default=2)
    ap.add_argument("--km-forms", default="fp8,bf16x3,fp64")
    ap.add_argument("--km-rows", type=int, default=10_000_000)
    ap.add_argument("--km-iters", type=int, default=50)
    ap.add_argument("--forms", default="bf16x3,fp64",
                    help="comma list; the first is reported as value (fp64, exact, bf16, bf16x3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--queries", type=int, default=100_000,
                    help="query count (100,000 is the benchmark; fewer only for profiling runs)")
    args = ap.parse_args()
    args.forms = [f for f in args.forms.split(",") if f]
    args.km_forms = [f for f in args.km_forms.split(",") if f]
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        copy_calibration()
        return
    if args.workload == "kmeans":
        print(json.dumps(run_kmeans(args)), flush=True)
        return
    if args.workload == "config1":
        print(json.dumps(run_config1(args)), flush=True)
        return
    if args.workload == "router":
        print(json.dumps(run_router(args)), flush=True)
        return
    if args.workload == "batched":
        print(json.dumps(run_batched(args)), flush=True)
        return
    line = run_knn(args)
    if args.workload == "all":
        line["kmeans"] = run_kmeans(args)
        line["kmeans_config1"] = run_config1(args)
        line["router"] = run_router(args)
        line["batched_kmeans"] = run_batched(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
