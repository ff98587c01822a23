"""KNeighborsClassifier / KNeighborsRegressor / NearestNeighbors with the
reference's builder API, and the k-neighbours graph.

Mirrors ferrolearn-neighbors/src/knn.rs: `with_n_neighbors / with_algorithm /
with_weights` (knn.rs:208-227), `fit(x, y)` (:256-320, :660-718), `predict`
(:322-383, :720-781), `predict_proba` (:452-486), `score` (:496-511,
:784-800), `kneighbors` (:823-834), `classes` (HasClasses) and
`n_samples_fit`; ferrolearn-neighbors/src/nearest_neighbors.rs
(`NearestNeighbors`, fit :198-228, kneighbors :255-302) and graph.rs
(`GraphMode` :26-32, `kneighbors_graph` :47-57 and the method forms
:225-304).  Every algorithm choice runs the GPU brute force: KdTree and
BallTree return the same neighbour sets (the reference's own cross-check,
nearest_neighbors.rs:657-700), so the answer does not depend on it.
"""
from __future__ import annotations

import copy
import ctypes as C
import enum

import numpy as np

from . import _lib as L
from .cluster import _dev2d, _dev_empty, _dtype_code, _host2d, _is_device, _to_host
from .errors import check


class Algorithm(enum.Enum):
    """knn.rs:58-68"""
    Auto = 0
    BruteForce = 1
    KdTree = 2
    BallTree = 3


class Weights(enum.Enum):
    """knn.rs:72-79"""
    Uniform = L.FL_UNIFORM
    Distance = L.FL_DISTANCE


class GraphMode(enum.Enum):
    """graph.rs:26-32: the value stored for every neighbour edge."""
    Connectivity = L.FL_GRAPH_CONNECTIVITY
    Distance = L.FL_GRAPH_DISTANCE


class CsrMatrix:
    """Compressed sparse rows, as ferrolearn_sparse::CsrMatrix: `indptr`
    (n_rows + 1), `indices` and `data` (nnz), numpy arrays or device tensors
    (a device query gives device arrays)."""

    def __init__(self, n_rows, n_cols, indptr, indices, data):
        self.shape = (int(n_rows), int(n_cols))
        self.indptr, self.indices, self.data = indptr, indices, data

    def n_rows(self) -> int:
        return self.shape[0]

    def n_cols(self) -> int:
        return self.shape[1]

    def nnz(self) -> int:
        return int(self.indices.shape[0])

    def _host(self):
        return tuple(_to_host(a) if _is_device(a) else a
                     for a in (self.indptr, self.indices, self.data))

    def to_dense(self) -> np.ndarray:
        indptr, indices, data = self._host()
        out = np.zeros(self.shape, np.float64)
        rows = np.repeat(np.arange(self.shape[0]), np.diff(indptr))
        out[rows, indices] = data
        return out

    def to_scipy(self):
        import scipy.sparse as sp
        indptr, indices, data = self._host()
        return sp.csr_matrix((data, indices, indptr), shape=self.shape)


class _Builder:
    def __init__(self):
        self.n_neighbors = 5
        self.algorithm = Algorithm.Auto
        self.weights = Weights.Uniform
        self.precision = "fp64"

    def _with(self, **kw):
        o = copy.copy(self)
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def with_n_neighbors(self, k: int):
        return self._with(n_neighbors=int(k))

    def with_algorithm(self, a: Algorithm):
        return self._with(algorithm=a)

    def with_weights(self, w: Weights):
        return self._with(weights=w)

    def with_precision(self, p: str):
        return self._with(precision=p)


class _Fitted:
    _is_classifier = True

    def __init__(self, handle, d, n_neighbors, precision):
        self._h = C.c_void_p(handle)
        self._d = d
        self.n_neighbors = n_neighbors
        self.precision = precision
        self.report = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().fl_knn_free(h)
            self._h = None

    def set_precision(self, p: str):
        check(L.lib().fl_knn_set_precision(self._h, L.PRECISIONS[p]))
        self.precision = p

    def n_samples_fit(self) -> int:
        return int(L.lib().fl_knn_n_samples_fit(self._h))

    def _query(self, x):
        if _is_device(x):
            return _dev2d(x), True
        return _host2d(x), False

    @staticmethod
    def _k_arg(n_neighbors):
        # None: the fitted n_neighbors (0 in the C call); an explicit 0 is
        # the reference's InvalidParameter (a negative k in the C call)
        if n_neighbors is None:
            return 0
        k = int(n_neighbors)
        return -1 if k == 0 else k

    def kneighbors(self, x, n_neighbors=None):
        """knn.rs:823-834: (distances, indices), each n_queries x k."""
        k = self._k_arg(n_neighbors)
        q, dev = self._query(x)
        nq, d = q.shape
        kk = self.n_neighbors if k == 0 else k
        rep = L.fl_report()
        if dev:
            dist = _dev_empty((nq, max(kk, 0)), np.float64, q)
            idx = _dev_empty((nq, max(kk, 0)), np.int64, q)
            check(L.lib().fl_knn_kneighbors_device(self._h, C.c_void_p(q.data_ptr()),
                                                   _dtype_code(q.dtype), nq, d, k,
                                                   C.c_void_p(dist.data_ptr()),
                                                   C.c_void_p(idx.data_ptr()), C.byref(rep)))
        else:
            dist = np.zeros((nq, max(kk, 0)), np.float64)
            idx = np.zeros((nq, max(kk, 0)), np.int64)
            check(L.lib().fl_knn_kneighbors(self._h, q.ctypes.data_as(C.c_void_p),
                                            _dtype_code(q.dtype), nq, d, k,
                                            dist.ctypes.data_as(C.c_void_p),
                                            idx.ctypes.data_as(C.c_void_p), C.byref(rep)))
        self.report = rep.as_dict()
        return dist, idx


    def kneighbors_graph(self, x, n_neighbors=None, mode=GraphMode.Connectivity) -> CsrMatrix:
        """graph.rs:225-304: the k-neighbours graph of the rows of x against
        the fitted rows, n_queries x n_samples_fit."""
        k = self._k_arg(n_neighbors)
        q, dev = self._query(x)
        nq, d = q.shape
        kk = max(self.n_neighbors if k == 0 else k, 0)
        rep = L.fl_report()
        if dev:
            indptr = _dev_empty((nq + 1,), np.int64, q)
            indices = _dev_empty((nq * kk,), np.int64, q)
            data = _dev_empty((nq * kk,), np.float64, q)
            check(L.lib().fl_knn_kneighbors_graph_device(
                self._h, C.c_void_p(q.data_ptr()), _dtype_code(q.dtype), nq, d, k, mode.value,
                C.c_void_p(indptr.data_ptr()), C.c_void_p(indices.data_ptr()),
                C.c_void_p(data.data_ptr()), C.byref(rep)))
        else:
            indptr = np.zeros(nq + 1, np.int64)
            indices = np.zeros(nq * kk, np.int64)
            data = np.zeros(nq * kk, np.float64)
            check(L.lib().fl_knn_kneighbors_graph(
                self._h, q.ctypes.data_as(C.c_void_p), _dtype_code(q.dtype), nq, d, k,
                mode.value, indptr.ctypes.data_as(C.c_void_p), indices.ctypes.data_as(C.c_void_p),
                data.ctypes.data_as(C.c_void_p), C.byref(rep)))
        self.report = rep.as_dict()
        return CsrMatrix(nq, self.n_samples_fit(), indptr, indices, data)


class FittedNearestNeighbors(_Fitted):
    """nearest_neighbors.rs:177-417 (kneighbors and the graph; no targets)."""
    _is_classifier = False

    def shape(self):
        return (self.n_samples_fit(), self._d)


class FittedKNeighborsClassifier(_Fitted):
    def classes(self):
        n = int(L.lib().fl_knn_n_classes(self._h))
        out = np.zeros(n, np.int64)
        check(L.lib().fl_knn_classes(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def n_classes(self) -> int:
        return int(L.lib().fl_knn_n_classes(self._h))

    def predict(self, x):
        q, dev = self._query(x)
        nq, d = q.shape
        rep = L.fl_report()
        if dev:
            out = _dev_empty((nq,), np.int64, q)
            check(L.lib().fl_knn_predict_classify_device(self._h, C.c_void_p(q.data_ptr()),
                                                         _dtype_code(q.dtype), nq, d,
                                                         C.c_void_p(out.data_ptr()),
                                                         C.byref(rep)))
        else:
            out = np.zeros(nq, np.int64)
            check(L.lib().fl_knn_predict_classify(self._h, q.ctypes.data_as(C.c_void_p),
                                                  _dtype_code(q.dtype), nq, d,
                                                  out.ctypes.data_as(C.c_void_p), C.byref(rep)))
        self.report = rep.as_dict()
        return out

    def predict_proba(self, x):
        q = _host2d(_to_host(x) if _is_device(x) else x)
        nq, d = q.shape
        out = np.zeros((nq, self.n_classes()), np.float64)
        rep = L.fl_report()
        check(L.lib().fl_knn_predict_proba(self._h, q.ctypes.data_as(C.c_void_p),
                                           _dtype_code(q.dtype), nq, d,
                                           out.ctypes.data_as(C.c_void_p), C.byref(rep)))
        self.report = rep.as_dict()
        return out

    def score(self, x, y):
        """knn.rs:496-511: mean accuracy."""
        y = np.asarray(y)
        if x.shape[0] != y.shape[0]:
            from .errors import ShapeMismatch
            raise ShapeMismatch("y length must match number of samples in X")
        p = self.predict(x)
        if _is_device(p):
            p = _to_host(p)
        return float((p == y).sum()) / y.shape[0] if y.shape[0] else 0.0


class FittedKNeighborsRegressor(_Fitted):
    _is_classifier = False

    def predict(self, x):
        q, dev = self._query(x)
        nq, d = q.shape
        rep = L.fl_report()
        if dev:
            out = _dev_empty((nq,), np.float64, q)
            check(L.lib().fl_knn_predict_regress_device(self._h, C.c_void_p(q.data_ptr()),
                                                        _dtype_code(q.dtype), nq, d,
                                                        C.c_void_p(out.data_ptr()),
                                                        C.byref(rep)))
        else:
            out = np.zeros(nq, np.float64)
            check(L.lib().fl_knn_predict_regress(self._h, q.ctypes.data_as(C.c_void_p),
                                                 _dtype_code(q.dtype), nq, d,
                                                 out.ctypes.data_as(C.c_void_p), C.byref(rep)))
        self.report = rep.as_dict()
        return out

    def score(self, x, y):
        """knn.rs:784-800 (R^2 via r2_score, knn.rs:846-870)."""
        y = np.asarray(y, dtype=np.float64)
        p = self.predict(x)
        if _is_device(p):
            p = _to_host(p)
        n = y.shape[0]
        if n == 0:
            return 0.0
        mean = 0.0
        for v in y:
            mean = mean + v
        mean = mean / n
        ss_res = ss_tot = 0.0
        for i in range(n):
            r = y[i] - p[i]
            t = y[i] - mean
            ss_res = ss_res + r * r
            ss_tot = ss_tot + t * t
        if ss_tot == 0.0:
            return 1.0 if ss_res == 0.0 else float("-inf")
        return 1.0 - ss_res / ss_tot


class KNeighborsClassifier(_Builder):
    """knn.rs:184-254.  Defaults: n_neighbors 5, Auto, Uniform."""

    def fit(self, x, y) -> FittedKNeighborsClassifier:
        xh = _host2d(_to_host(x) if _is_device(x) else x)
        yh = np.ascontiguousarray(np.asarray(y, dtype=np.int64))
        n, d = xh.shape
        h = C.c_void_p()
        check(L.lib().fl_knn_fit_classifier(xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype),
                                            n, d, yh.ctypes.data_as(C.c_void_p), yh.shape[0],
                                            self.n_neighbors, self.weights.value,
                                            L.PRECISIONS[self.precision], C.byref(h)))
        return FittedKNeighborsClassifier(h.value, d, self.n_neighbors, self.precision)


class KNeighborsRegressor(_Builder):
    """knn.rs:590-658.  Defaults: n_neighbors 5, Auto, Uniform."""

    def fit(self, x, y) -> FittedKNeighborsRegressor:
        xh = _host2d(_to_host(x) if _is_device(x) else x)
        yh = np.ascontiguousarray(np.asarray(y, dtype=np.float64))
        n, d = xh.shape
        h = C.c_void_p()
        check(L.lib().fl_knn_fit_regressor(xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype),
                                           n, d, yh.ctypes.data_as(C.c_void_p), yh.shape[0],
                                           self.n_neighbors, self.weights.value,
                                           L.PRECISIONS[self.precision], C.byref(h)))
        return FittedKNeighborsRegressor(h.value, d, self.n_neighbors, self.precision)


class NearestNeighbors(_Builder):
    """nearest_neighbors.rs:64-128.  Defaults: n_neighbors 5, Auto, leaf_size
    30 (kept for the builder; the brute force has no leaves)."""

    def __init__(self):
        super().__init__()
        self.leaf_size = 30

    def with_leaf_size(self, leaf_size: int):
        return self._with(leaf_size=int(leaf_size))

    def fit(self, x, y=None) -> FittedNearestNeighbors:
        xh = _host2d(_to_host(x) if _is_device(x) else x)
        n, d = xh.shape
        h = C.c_void_p()
        check(L.lib().fl_nn_fit(xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype), n, d,
                                self.n_neighbors, L.PRECISIONS[self.precision], C.byref(h)))
        return FittedNearestNeighbors(h.value, d, self.n_neighbors, self.precision)


def kneighbors_graph(x, n_neighbors, mode=GraphMode.Connectivity, precision="fp64") -> CsrMatrix:
    """graph.rs:47-57: NearestNeighbors over x (Auto), then the graph of x
    against itself, n_samples x n_samples."""
    nn = NearestNeighbors().with_n_neighbors(n_neighbors).with_precision(precision).fit(x)
    return nn.kneighbors_graph(x, n_neighbors, mode)
