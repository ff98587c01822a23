"""Tile-shape autotuner of the batched KMeans assignment kernel.

fl_kmeans_fit_batched scores each row exactly against the centres of every
restart of its problem.  Its kernel has several tile shapes ("variants"):
D feature registers per row (d <= D), U centres folded side by side, and
R rows per lane.  The variant is chosen per (d bucket, k bucket) from a
table of 4 x 4 entries (d <= 8, 16, 32, 64; k <= 16, 64, 256, 1024).

`autotune()` times every variant of each bucket on a synthetic problem set
with the batched call itself, and returns the fastest per bucket.  `save()`
writes the table as JSON.  The library loads `tuning_gfx950.json` from this
folder the first time it is opened.  Results never depend on the table; only
speed does.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

from . import _lib as L
from .errors import check

D_BUCKETS = (8, 16, 32, 64)
K_BUCKETS = (16, 64, 256, 1024)
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")


def _bucket(values, x):
    for i, v in enumerate(values):
        if x <= v:
            return i
    raise ValueError(f"{x} is beyond the batched kernel's buckets {values}")


def variants():
    """[(v, D, U, R)] for every variant of the kernel."""
    lib = L.lib()
    out = []
    for v in range(int(lib.fl_kmeans_batched_n_variants())):
        D, U, R = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib.fl_kmeans_batched_variant(v, C.byref(D), C.byref(U), C.byref(R)))
        out.append((v, D.value, U.value, R.value))
    return out


def variants_for(d):
    """Variant indices usable for d features (those with d's bucket D)."""
    D = D_BUCKETS[_bucket(D_BUCKETS, d)]
    return [v for v, vd, _, _ in variants() if vd == D]


def get_raw_table():
    t = (C.c_int32 * 16)()
    check(L.lib().fl_kmeans_batched_get_tuning(t))
    return [int(x) for x in t]


def set_raw_table(values):
    t = (C.c_int32 * 16)(*[int(x) for x in values])
    check(L.lib().fl_kmeans_batched_set_tuning(t))


def get_table():
    """{(D, k bucket upper bound): variant}."""
    raw = get_raw_table()
    return {(D, K): raw[4 * i + j] for i, D in enumerate(D_BUCKETS)
            for j, K in enumerate(K_BUCKETS)}


def set_table(entries):
    """Set the entries {(d, k): variant} for the buckets of each (d, k)."""
    raw = get_raw_table()
    for (d, k), v in entries.items():
        raw[4 * _bucket(D_BUCKETS, d) + _bucket(K_BUCKETS, k)] = int(v)
    set_raw_table(raw)


def save(table, path=DEFAULT_PATH, meta=None):
    doc = {"table": {f"{D}x{K}": int(v) for (D, K), v in table.items()},
           "variants": {str(v): {"D": D, "U": U, "R": R} for v, D, U, R in variants()}}
    if meta:
        doc["meta"] = meta
    with open(path, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)


def load(path=DEFAULT_PATH):
    """Load a saved table into the library (returns it)."""
    with open(path) as f:
        doc = json.load(f)
    raw = get_raw_table()
    for key, v in doc["table"].items():
        D, K = (int(x) for x in key.split("x"))
        raw[4 * D_BUCKETS.index(D) + K_BUCKETS.index(K)] = int(v)
    set_raw_table(raw)
    return get_table()


def _bucket_problems(D, K, n_problems, n_rows, seed):
    """Synthetic problems of one bucket: d = D, k near the bucket's upper end,
    Gaussian blobs, random rows as the initial centres of 8 restarts."""
    from .device import to_device
    from .cluster import KMeans
    g = np.random.default_rng(seed)
    k = max(2, (K * 3) // 4)
    probs = []
    for _ in range(n_problems):
        cen = g.uniform(-10, 10, size=(k, D))
        X = (cen[g.integers(0, k, n_rows)] + g.normal(size=(n_rows, D))).astype(np.float32)
        init = X[g.integers(0, n_rows, size=8 * k)].astype(np.float64).reshape(8, k, D)
        m = KMeans(k).with_n_init(8).with_max_iter(4).with_tol(0.0)
        probs.append((to_device(X), m, init))
    return probs


def autotune(n_problems=8, n_rows=20000, reps=3, seed=0, log=None):
    """Time every variant per bucket; returns (table, timings).  The Lloyd
    phase of a batched call (its device time) is the measure."""
    from .cluster import kmeans_fit_batched
    raw0 = get_raw_table()
    table, timings = {}, {}
    try:
        for D in D_BUCKETS:
            for K in K_BUCKETS:
                probs = _bucket_problems(D, K, n_problems, n_rows, seed)
                best = None
                for v in variants_for(D):
                    set_table({(D, K): v})
                    kmeans_fit_batched(probs)            # warm-up
                    ms = []
                    for _ in range(reps):
                        t0 = time.perf_counter()
                        _, rep = kmeans_fit_batched(probs)
                        ms.append(rep["ms_distance"] if rep["ms_distance"] > 0
                                  else (time.perf_counter() - t0) * 1e3)
                    t = float(np.median(ms))
                    timings[f"{D}x{K}:{v}"] = t
                    if log:
                        log(f"tune D={D} k<={K} variant {v}: {t:.3f} ms")
                    if best is None or t < best[1]:
                        best = (v, t)
                table[(D, K)] = best[0]
                del probs
    finally:
        set_raw_table(raw0)
    return table, timings
