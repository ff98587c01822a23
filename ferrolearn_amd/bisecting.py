"""BisectingKMeans on the GPU (ferrolearn-cluster/src/bisecting_kmeans.rs).

The reference's loop with its StdRng seeds; the assignments, the centre sums
(in member order) and the SSE folds run on the device.  Labels, centres and
inertia are the reference's bit for bit.
"""
from __future__ import annotations

import copy
import ctypes as C
import enum

import numpy as np

from . import _lib as L
from .cluster import FittedKMeans, _dev2d, _dev_empty, _dtype_code, _host2d, _is_device
from .errors import ShapeMismatch, check


class BisectingStrategy(enum.Enum):
    """bisecting_kmeans.rs:54-61."""
    LargestCluster = 0
    LargestSSE = 1


class BisectingKMeans:
    """bisecting_kmeans.rs:72-130.  Defaults from BisectingKMeans::new:
    max_iter 300, n_init 10, random_state None (seed 0), LargestCluster."""

    def __init__(self, n_clusters: int):
        self.n_clusters = int(n_clusters)
        self.max_iter = 300
        self.n_init = 10
        self.random_state = None
        self.bisecting_strategy = BisectingStrategy.LargestCluster

    def _with(self, **kw):
        o = copy.copy(self)
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def with_max_iter(self, max_iter: int):
        return self._with(max_iter=int(max_iter))

    def with_n_init(self, n_init: int):
        return self._with(n_init=int(n_init))

    def with_random_state(self, seed: int):
        return self._with(random_state=int(seed))

    def with_bisecting_strategy(self, strategy: BisectingStrategy):
        return self._with(bisecting_strategy=BisectingStrategy(strategy))

    def _params(self):
        p = L.fl_bisecting_params()
        L.lib().fl_bisecting_params_default(C.byref(p), self.n_clusters)
        p.max_iter = self.max_iter
        p.n_init = self.n_init
        if self.random_state is not None:
            p.random_state = self.random_state & (2 ** 64 - 1)
            p.has_random_state = 1
        p.strategy = self.bisecting_strategy.value
        return p

    def fit(self, x, y=None) -> "FittedBisectingKMeans":
        """Fit<Array2<F>, ()> (bisecting_kmeans.rs:326-505)."""
        p = self._params()
        inertia = C.c_double()
        nc = C.c_int64()
        if _is_device(x):
            x = _dev2d(x)
            n, d = x.shape
            centers = _dev_empty((max(self.n_clusters, 0), d), np.float64, x)
            labels = _dev_empty((n,), np.int64, x)
            check(L.lib().fl_bisecting_kmeans_fit_device(
                C.c_void_p(x.data_ptr()), _dtype_code(x.dtype), n, d, C.byref(p),
                C.c_void_p(centers.data_ptr()), C.c_void_p(labels.data_ptr()), C.byref(inertia),
                C.byref(nc)))
            return FittedBisectingKMeans(centers[:nc.value], labels, inertia.value)
        xh = _host2d(x)
        n, d = xh.shape
        centers = np.zeros((max(self.n_clusters, 0), d), np.float64)
        labels = np.zeros(n, np.int64)
        check(L.lib().fl_bisecting_kmeans_fit(
            xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype), n, d, C.byref(p),
            centers.ctypes.data_as(C.c_void_p), labels.ctypes.data_as(C.c_void_p),
            C.byref(inertia), C.byref(nc)))
        return FittedBisectingKMeans(centers[:nc.value].copy(), labels, inertia.value)

    def fit_predict(self, x):
        """bisecting_kmeans.rs:515-518: fit, then predict on the same data."""
        return self.fit(x).predict(x)


class FittedBisectingKMeans:
    """bisecting_kmeans.rs:138-171, with Predict (:531-570) and Transform
    (:581-605), which compute what KMeans' predict and transform do."""

    def __init__(self, centers, labels, inertia):
        self.cluster_centers_ = centers
        self.labels_ = labels
        self.inertia_ = float(inertia)
        self._km = FittedKMeans(centers, labels, inertia, 0, {})

    def cluster_centers(self):
        return self.cluster_centers_

    def labels(self):
        return self.labels_

    def inertia(self) -> float:
        return self.inertia_

    def n_clusters(self) -> int:
        return int(self.cluster_centers_.shape[0])

    def _check(self, x, what):
        d = x.shape[1] if len(x.shape) == 2 else -1
        want = self.cluster_centers_.shape[1]
        if d != want:
            raise ShapeMismatch(f"ShapeMismatch: expected [{want}], actual [{d}]: number of "
                                f"features must match fitted BisectingKMeans{what}")

    def predict(self, x):
        self._check(x, " model")
        return self._km.predict(x)

    def transform(self, x):
        self._check(x, "")
        return self._km.transform(x)
