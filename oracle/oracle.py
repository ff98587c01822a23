"""ctypes wrapper over oracle/build/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of ferrolearn's f64 squared-Euclidean
path (ref_cpu.c).  Only tests/, the smoke check and bench.py's CPU-baseline
leg may import this module; the product never does.

Every function takes numpy arrays; f32 inputs are widened to f64, which is
exact, exactly as the reference's f64 path sees f32 data converted to f64.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


def build() -> None:
    """Compile ref_cpu.c with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_sqeuclid.restype = C.c_double
        L.or_sqeuclid.argtypes = [_dp, _dp, _i64]
        L.or_euclid.restype = C.c_double
        L.or_euclid.argtypes = [_dp, _dp, _i64]
        L.or_nearest_center.restype = _i64
        L.or_nearest_center.argtypes = [_dp, _dp, _i64, _i64, _dp]
        L.or_assign.restype = C.c_double
        L.or_assign.argtypes = [_dp, _i64, _i64, _dp, _i64, _ip, C.c_int]
        L.or_assign_rows.restype = None
        L.or_assign_rows.argtypes = [_dp, _i64, _i64, _dp, _i64, _ip, _dp, C.c_int]
        L.or_recompute_centroids.restype = C.c_double
        L.or_recompute_centroids.argtypes = [_dp, _i64, _i64, _ip, _dp, _i64, _dp, _dp]
        L.or_kmeans_lloyd.restype = None
        L.or_kmeans_lloyd.argtypes = [_dp, _i64, _i64, _i64, _dp, _i64, C.c_double,
                                      C.c_int, _dp, _ip, _dp, _ip]
        L.or_kmeans_fit_from.restype = _i64
        L.or_kmeans_fit_from.argtypes = [_dp, _i64, _i64, _i64, _dp, _i64, _i64,
                                         C.c_double, C.c_int, _dp, _ip, _dp, _ip]
        L.or_transform.restype = None
        L.or_transform.argtypes = [_dp, _i64, _i64, _dp, _i64, _dp, C.c_int]
        L.or_rng_stream.restype = None
        L.or_rng_stream.argtypes = [C.c_uint64, _i64, C.POINTER(C.c_uint64)]
        L.or_blocked_sum.restype = C.c_double
        L.or_blocked_sum.argtypes = [_dp, _i64]
        L.or_blocked_search.restype = _i64
        L.or_blocked_search.argtypes = [_dp, _i64, C.c_double]
        L.or_kmeanspp.restype = None
        L.or_kmeanspp.argtypes = [_dp, _i64, _i64, _i64, C.c_uint64, _dp, C.c_int]
        L.or_minibatch_update.restype = C.c_double
        L.or_minibatch_update.argtypes = [_dp, _i64, _ip, _ip, _i64, _dp, _i64, _ip]
        L.or_minibatch_run.restype = None
        L.or_minibatch_run.argtypes = [_dp, _i64, _i64, _i64, _dp, _i64, _i64,
                                       C.c_double, C.c_uint64, C.c_int, _dp, _ip, _dp, _ip]
        L.or_minibatch_fit.restype = None
        L.or_minibatch_fit.argtypes = [_dp, _i64, _i64, _i64, _i64, _i64, C.c_double, _i64,
                                       C.c_uint64, C.c_int, _dp, _ip, _dp, _ip]
        L.or_knn_query.restype = _i64
        L.or_knn_query.argtypes = [_dp, _i64, _i64, _dp, _i64, _ip, _dp]
        L.or_kneighbors.restype = None
        L.or_kneighbors.argtypes = [_dp, _i64, _i64, _dp, _i64, _i64, _ip, _dp, C.c_int]
        _u32p = C.POINTER(C.c_uint32)
        L.or_chacha_block.restype = None
        L.or_chacha_block.argtypes = [_u32p, C.c_int, _u32p]
        L.or_stdrng_stream.restype = None
        L.or_stdrng_stream.argtypes = [C.c_uint64, _i64, _u32p]
        L.or_stdrng_ranges.restype = None
        L.or_stdrng_ranges.argtypes = [C.c_uint64, _ip, _ip, _i64, _ip]
        L.or_stdrng_f64s.restype = None
        L.or_stdrng_f64s.argtypes = [C.c_uint64, _i64, _dp]
        L.or_kmeanspp_stdrng.restype = None
        L.or_kmeanspp_stdrng.argtypes = [_dp, _i64, _i64, _i64, C.c_uint64, _dp, C.c_int]
        L.or_kmeans_fit_stdrng.restype = None
        L.or_kmeans_fit_stdrng.argtypes = [_dp, _i64, _i64, _i64, _i64, C.c_double, _i64,
                                           C.c_uint64, C.c_int, _dp, _ip, _dp, _ip]
        L.or_minibatch_fit_stdrng.restype = None
        L.or_minibatch_fit_stdrng.argtypes = [_dp, _i64, _i64, _i64, _i64, _i64, C.c_double,
                                              _i64, C.c_uint64, C.c_int, _dp, _ip, _dp, _ip]
        L.or_bisecting_kmeans.restype = _i64
        L.or_bisecting_kmeans.argtypes = [_dp, _i64, _i64, _i64, _i64, _i64, C.c_uint64, C.c_int,
                                          _i64, _dp, _ip, _dp]
        L.or_euclidean_distances.restype = None
        L.or_euclidean_distances.argtypes = [_dp, _i64, _dp, _i64, _i64, _dp, C.c_int]
        L.or_pairwise_argmin_min.restype = None
        L.or_pairwise_argmin_min.argtypes = [_dp, _i64, _dp, _i64, _i64, _ip, _dp, C.c_int]
        L.or_class_scores.restype = None
        L.or_class_scores.argtypes = [_ip, _dp, _i64, _ip, _i64, C.c_int, _dp]
        L.or_vote.restype = _i64
        L.or_vote.argtypes = [_dp, _i64]
        L.or_proba.restype = None
        L.or_proba.argtypes = [_dp, _i64, _dp]
        L.or_weighted_mean.restype = C.c_double
        L.or_weighted_mean.argtypes = [_ip, _dp, _i64, _dp, C.c_int]
        L.or_knn_classify.restype = None
        L.or_knn_classify.argtypes = [_dp, _i64, _i64, _ip, _i64, _dp, _i64, _i64,
                                      C.c_int, _ip, _dp, C.c_int]
        L.or_knn_regress.restype = None
        L.or_knn_regress.argtypes = [_dp, _i64, _i64, _dp, _dp, _i64, _i64, C.c_int,
                                     _dp, C.c_int]
        L.or_max_threads.restype = C.c_int
        L.or_max_threads.argtypes = []
        _lib = L
    return _lib


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i64a(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _p(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def max_threads() -> int:
    return int(lib().or_max_threads())


def sqeuclid(a, b) -> float:
    a, b = _f64(a), _f64(b)
    return float(lib().or_sqeuclid(_p(a), _p(b), a.shape[0]))


def euclid(a, b) -> float:
    a, b = _f64(a), _f64(b)
    return float(lib().or_euclid(_p(a), _p(b), a.shape[0]))


def nearest_center(row, centers):
    row, centers = _f64(row), _f64(centers)
    dist = C.c_double()
    lab = lib().or_nearest_center(_p(row), _p(centers), centers.shape[0],
                                  centers.shape[1], C.byref(dist))
    return int(lab), float(dist.value)


def assign(X, centers, nthreads=0):
    X, centers = _f64(X), _f64(centers)
    n, d = X.shape
    labels = np.zeros(n, np.int64)
    inertia = lib().or_assign(_p(X), n, d, _p(centers), centers.shape[0], _pi(labels), nthreads)
    return labels, float(inertia)


def assign_rows(X, centers, nthreads=0):
    X, centers = _f64(X), _f64(centers)
    n, d = X.shape
    labels = np.zeros(n, np.int64)
    dist = np.zeros(n, np.float64)
    lib().or_assign_rows(_p(X), n, d, _p(centers), centers.shape[0], _pi(labels), _p(dist),
                         nthreads)
    return labels, dist


def recompute_centroids(X, labels, old_centers):
    X, labels, old = _f64(X), _i64a(labels), _f64(old_centers)
    n, d = X.shape
    k = old.shape[0]
    new = np.zeros((k, d), np.float64)
    counts = np.zeros(k, np.float64)
    shift = lib().or_recompute_centroids(_p(X), n, d, _pi(labels), _p(old), k, _p(new),
                                         _p(counts))
    return new, counts, float(shift)


def kmeans_lloyd(X, init, max_iter, tol, nthreads=0):
    X, init = _f64(X), _f64(init)
    n, d = X.shape
    k = init.shape[0]
    centers = np.zeros((k, d), np.float64)
    labels = np.zeros(n, np.int64)
    inertia = C.c_double()
    n_iter = C.c_int64()
    lib().or_kmeans_lloyd(_p(X), n, d, k, _p(init), max_iter, tol, nthreads, _p(centers),
                          _pi(labels), C.byref(inertia), C.byref(n_iter))
    return centers, labels, float(inertia.value), int(n_iter.value)


def kmeans_fit_from(X, inits, max_iter, tol, nthreads=0):
    """inits: (n_init, k, d)."""
    X, inits = _f64(X), _f64(inits)
    n, d = X.shape
    n_init, k, _ = inits.shape
    centers = np.zeros((k, d), np.float64)
    labels = np.zeros(n, np.int64)
    inertia = C.c_double()
    n_iter = C.c_int64()
    best = lib().or_kmeans_fit_from(_p(X), n, d, k, _p(inits), n_init, max_iter, tol,
                                    nthreads, _p(centers), _pi(labels), C.byref(inertia),
                                    C.byref(n_iter))
    return centers, labels, float(inertia.value), int(n_iter.value), int(best)


def transform(X, centers, nthreads=0):
    X, centers = _f64(X), _f64(centers)
    n, d = X.shape
    k = centers.shape[0]
    out = np.zeros((n, k), np.float64)
    lib().or_transform(_p(X), n, d, _p(centers), k, _p(out), nthreads)
    return out


def rng_stream(seed, count):
    out = np.zeros(count, np.uint64)
    lib().or_rng_stream(C.c_uint64(seed), count, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def blocked_sum(v) -> float:
    v = _f64(v)
    return float(lib().or_blocked_sum(_p(v), v.shape[0]))


def blocked_search(v, threshold) -> int:
    v = _f64(v)
    return int(lib().or_blocked_search(_p(v), v.shape[0], threshold))


def kmeanspp(X, k, seed, nthreads=0):
    X = _f64(X)
    n, d = X.shape
    out = np.zeros((k, d), np.float64)
    lib().or_kmeanspp(_p(X), n, d, k, C.c_uint64(seed), _p(out), nthreads)
    return out


def minibatch_update(X, batch_idx, batch_lab, centers, counts):
    X = _f64(X)
    bi, bl = _i64a(batch_idx), _i64a(batch_lab)
    C_ = _f64(centers).copy()
    cnt = _i64a(counts).copy()
    shift = lib().or_minibatch_update(_p(X), X.shape[1], _pi(bi), _pi(bl), bi.shape[0], _p(C_),
                                      C_.shape[0], _pi(cnt))
    return C_, cnt, float(shift)


def minibatch_run(X, init, batch, max_iter, tol, seed, nthreads=0):
    X, init = _f64(X), _f64(init)
    n, d = X.shape
    k = init.shape[0]
    centers = np.zeros((k, d), np.float64)
    labels = np.zeros(n, np.int64)
    inertia = C.c_double()
    n_iter = C.c_int64()
    lib().or_minibatch_run(_p(X), n, d, k, _p(init), batch, max_iter, tol, C.c_uint64(seed),
                           nthreads, _p(centers), _pi(labels), C.byref(inertia),
                           C.byref(n_iter))
    return centers, labels, float(inertia.value), int(n_iter.value)


def minibatch_fit(X, k, batch, max_iter, tol, n_init, seed, nthreads=0):
    """A whole MiniBatchKMeans fit with the library's k-means++ per run."""
    X = _f64(X)
    n, d = X.shape
    centers = np.zeros((k, d), np.float64)
    labels = np.zeros(n, np.int64)
    inertia = C.c_double()
    n_iter = C.c_int64()
    lib().or_minibatch_fit(_p(X), n, d, k, batch, max_iter, C.c_double(tol), n_init,
                           C.c_uint64(seed), nthreads, _p(centers), _pi(labels),
                           C.byref(inertia), C.byref(n_iter))
    return centers, labels, float(inertia.value), int(n_iter.value)


def knn_query(data, q, k):
    data, q = _f64(data), _f64(q)
    N, d = data.shape
    kk = min(k, N)
    idx = np.zeros(max(kk, 1), np.int64)
    dist = np.zeros(max(kk, 1), np.float64)
    got = lib().or_knn_query(_p(data), N, d, _p(q), k, _pi(idx), _p(dist))
    return idx[:got], dist[:got]


def kneighbors(data, Q, k, nthreads=0):
    data, Q = _f64(data), _f64(Q)
    N, d = data.shape
    nq = Q.shape[0]
    idx = np.zeros((nq, k), np.int64)
    dist = np.zeros((nq, k), np.float64)
    lib().or_kneighbors(_p(data), N, d, _p(Q), nq, k, _pi(idx), _p(dist), nthreads)
    return dist, idx


def knn_to_csr(dist, idx, n_cols, distance_mode):
    """graph.rs:91-123 (knn_to_csr): per row, the (column, value) pairs of
    the k neighbours -- value 1.0 for connectivity, the distance otherwise --
    sorted by column (sort_by_key, stable), then concatenated.  Returns
    (indptr, indices, data, shape)."""
    nq, k = idx.shape
    indptr = np.arange(nq + 1, dtype=np.int64) * k
    indices = np.zeros(nq * k, np.int64)
    data = np.zeros(nq * k, np.float64)
    for i in range(nq):
        pairs = sorted(((int(idx[i, j]), float(dist[i, j]) if distance_mode else 1.0)
                        for j in range(k)), key=lambda p: p[0])
        for j, (c, v) in enumerate(pairs):
            indices[i * k + j] = c
            data[i * k + j] = v
    return indptr, indices, data, (nq, int(n_cols))


def bisecting_kmeans(X, n_clusters, max_iter=300, n_init=10, seed=0, strategy="size",
                     max_stall=1000):
    """bisecting_kmeans.rs:326-505: (centres, labels, inertia).  strategy
    "size" = LargestCluster, "sse" = LargestSSE.  Raises where the reference
    panics (no 2-means run below f64::MAX) or would keep drawing (max_stall
    bisections in a row with an empty side)."""
    X = _f64(X)
    n, d = X.shape
    cen = np.zeros((n_clusters, d), np.float64)
    lab = np.zeros(n, np.int64)
    inertia = np.zeros(1, np.float64)
    nc = lib().or_bisecting_kmeans(_p(X), n, d, n_clusters, max_iter, n_init, seed & (2**64 - 1),
                                   0 if strategy == "size" else 1, max_stall, _p(cen), _pi(lab),
                                   _p(inertia))
    if nc < 0:
        raise RuntimeError("bisecting_kmeans: " + ("no 2-means run below f64::MAX" if nc == -1
                                                   else "bisections keep leaving a side empty"))
    return cen[:nc].copy(), lab, float(inertia[0])


def euclidean_distances(X, Y, nthreads=0):
    """pairwise.rs:157-189: the (n, m) matrix of distances between the rows
    of X and Y in the expansion arithmetic."""
    X, Y = _f64(X), _f64(Y)
    n, d = X.shape
    m = Y.shape[0]
    out = np.zeros((n, m), np.float64)
    lib().or_euclidean_distances(_p(X), n, _p(Y), m, d, _p(out), nthreads)
    return out


def pairwise_argmin_min(X, Y, nthreads=0):
    """pairwise.rs:564-629 with Metric::Euclidean: (indices, minimum
    distances) of each row of X over the rows of Y."""
    X, Y = _f64(X), _f64(Y)
    n, d = X.shape
    m = Y.shape[0]
    idx = np.zeros(n, np.int64)
    mins = np.zeros(n, np.float64)
    lib().or_pairwise_argmin_min(_p(X), n, _p(Y), m, d, _pi(idx), _p(mins), nthreads)
    return idx, mins


# ------------------------------------------------------- reference stream --
# StdRng of rand 0.9.4 (ChaCha12) as restated in ref_cpu.c.

def chacha_block(state, rounds):
    """One ChaCha block (state: 16 u32 words)."""
    a = np.ascontiguousarray(np.asarray(state, dtype=np.uint32))
    out = np.zeros(16, np.uint32)
    lib().or_chacha_block(a.ctypes.data_as(C.POINTER(C.c_uint32)), rounds,
                          out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def stdrng_stream(seed, count):
    out = np.zeros(count, np.uint32)
    lib().or_stdrng_stream(C.c_uint64(seed), count, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def stdrng_ranges(seed, lo, hi):
    """random_range(lo[i]..hi[i]) for i in order, from one generator."""
    lo = _i64a(lo)
    hi = _i64a(hi)
    out = np.zeros(lo.shape[0], np.int64)
    lib().or_stdrng_ranges(C.c_uint64(seed), _pi(lo), _pi(hi), lo.shape[0], _pi(out))
    return out


def stdrng_f64s(seed, count):
    out = np.zeros(count, np.float64)
    lib().or_stdrng_f64s(C.c_uint64(seed), count, _p(out))
    return out


def kmeanspp_stdrng(X, k, seed, nthreads=0):
    X = _f64(X)
    n, d = X.shape
    out = np.zeros((k, d), np.float64)
    lib().or_kmeanspp_stdrng(_p(X), n, d, k, C.c_uint64(seed), _p(out), nthreads)
    return out


def kmeans_fit_stdrng(X, k, max_iter, tol, n_init, seed, nthreads=0):
    X = _f64(X)
    n, d = X.shape
    centers = np.zeros((k, d), np.float64)
    labels = np.zeros(n, np.int64)
    inertia = C.c_double()
    n_iter = C.c_int64()
    lib().or_kmeans_fit_stdrng(_p(X), n, d, k, max_iter, C.c_double(tol), n_init,
                               C.c_uint64(seed), nthreads, _p(centers), _pi(labels),
                               C.byref(inertia), C.byref(n_iter))
    return centers, labels, float(inertia.value), int(n_iter.value)


def minibatch_fit_stdrng(X, k, batch, max_iter, tol, n_init, seed, nthreads=0):
    X = _f64(X)
    n, d = X.shape
    centers = np.zeros((k, d), np.float64)
    labels = np.zeros(n, np.int64)
    inertia = C.c_double()
    n_iter = C.c_int64()
    lib().or_minibatch_fit_stdrng(_p(X), n, d, k, batch, max_iter, C.c_double(tol), n_init,
                                  C.c_uint64(seed), nthreads, _p(centers), _pi(labels),
                                  C.byref(inertia), C.byref(n_iter))
    return centers, labels, float(inertia.value), int(n_iter.value)


def class_positions(y):
    """Sorted unique classes and the position of each label among them
    (knn.rs:304-306: sort_unstable + dedup)."""
    y = _i64a(y)
    classes = np.unique(y)
    return classes, np.searchsorted(classes, y).astype(np.int64)


def knn_classify(data, y, Q, k, weights=0, nthreads=0, proba=False):
    data, Q = _f64(data), _f64(Q)
    classes, ypos = class_positions(y)
    N, d = data.shape
    nq = Q.shape[0]
    pred = np.zeros(nq, np.int64)
    pr = np.zeros((nq, classes.shape[0]), np.float64) if proba else None
    lib().or_knn_classify(_p(data), N, d, _pi(ypos), classes.shape[0], _p(Q), nq, k, weights,
                          _pi(pred), _p(pr) if proba else None, nthreads)
    if proba:
        return classes[pred], pr
    return classes[pred]


def knn_regress(data, y, Q, k, weights=0, nthreads=0):
    data, Q, y = _f64(data), _f64(Q), _f64(y)
    N, d = data.shape
    nq = Q.shape[0]
    pred = np.zeros(nq, np.float64)
    lib().or_knn_regress(_p(data), N, d, _p(y), _p(Q), nq, k, weights, _p(pred), nthreads)
    return pred


# ---------------------------------------------------------------- router net --
# The router network of the cell router, restated in numpy float64 from
# ferrolearn-neural/src/mlp.rs: forward z = a @ W + b with ReLU hidden layers
# (:217-261, :69-74), softmax rows (:106-127), loss -sum(y ln max(p, 1e-15)) /
# batch with delta p - y (:1326-1340), grad_w = a^T delta / batch + alpha W,
# grad_b = mean(delta), delta_prev = (delta @ W^T) * (z > 0) (:302-345).
# The soft target y of a row is the fraction of its valid (>= 0) target cells
# equal to each cell; a row without targets contributes nothing.  With emulate_bf16 the inputs, weights, hidden activations and
# deltas are rounded to bfloat16 where the GPU stores them in bfloat16, so
# that the comparison isolates the accumulation order.

def bf16_round(a) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def mlp_split(p, d, hidden, n_cells):
    shapes = [(d, hidden), (hidden,), (hidden, hidden), (hidden,), (hidden, n_cells),
              (n_cells,)]
    out, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], np.float64).reshape(s))
        o += n
    return out


def mlp_loss_grad(params, X, cells, d, hidden, n_cells, alpha=0.0, emulate_bf16=False):
    """(mean loss, flat gradient) of the router network on one batch."""
    r = bf16_round if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = mlp_split(params, d, hidden, n_cells)
    B, k = cells.shape
    x = r(X)
    w1, w2, w3 = r(W1), r(W2), r(W3)
    a1 = r(np.maximum(x @ w1 + b1, 0.0))
    a2 = r(np.maximum(a1 @ w2 + b2, 0.0))
    z3 = a2 @ w3 + b3
    z3 = z3 - z3.max(axis=1, keepdims=True)
    e = np.exp(z3)
    p = e / e.sum(axis=1, keepdims=True)
    y = np.zeros_like(p)
    for i in range(B):
        valid = [c for c in cells[i] if c >= 0]
        for c in valid:
            y[i, c] += 1.0 / len(valid)
    has = (y.sum(axis=1) > 0)[:, None]
    loss = float(-(y * np.log(np.maximum(p, 1e-15))).sum() / B)
    d3 = r(np.where(has, p - y, 0.0))
    gW3 = a2.T @ d3 / B + alpha * W3
    gb3 = d3.sum(axis=0) / B
    d2 = r((d3 @ w3.T) * (a2 > 0))
    gW2 = a1.T @ d2 / B + alpha * W2
    gb2 = d2.sum(axis=0) / B
    d1 = r((d2 @ w2.T) * (a1 > 0))
    gW1 = x.T @ d1 / B + alpha * W1
    gb1 = d1.sum(axis=0) / B
    grad = np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, grad
