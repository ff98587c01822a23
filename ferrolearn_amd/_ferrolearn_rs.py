"""The PyO3 classes of ferrolearn-python on this path, for its sklearn wrappers.

Mirrors `_RsKMeans` (ferrolearn-python/src/clusterers.rs:12-130),
`_RsKNeighborsClassifier` (classifiers.rs:298-353), `_RsKNeighborsRegressor`
(extras.rs:17-60, 381-389) and `_RsMiniBatchKMeans` (extras.rs:693-735):
the same constructor keywords and defaults, `fit` / `predict` / `transform`,
the fitted getters, float64 inputs, int64 label outputs, `ValueError` with
the reference's error text when a model call fails (the PyO3 mapping of
FerroError's Display, ferrolearn-core/src/error.rs:34-89) and
`RuntimeError("not fitted")` before fit.  The clustering classes draw from
the reference's StdRng stream (KMeans.with_rng("stdrng")).  The sklearn wrappers of
ferrolearn-python/python/ferrolearn import these names from
`ferrolearn._ferrolearn_rs`; binding that module name to this one runs them
on the GPU.
"""
from __future__ import annotations

import re

import numpy as np

from .cluster import KMeans, MiniBatchKMeans
from .errors import FerroError
from .neighbors import KNeighborsClassifier, KNeighborsRegressor

_PATTERNS = (
    (re.compile(r"^InvalidParameter: ([^:]+): (.*)$", re.S), "Invalid parameter `{0}`: {1}"),
    (re.compile(r"^InsufficientSamples: required (\d+), actual (\d+): (.*)$", re.S),
     "Insufficient samples: need at least {0}, got {1} ({2})"),
    (re.compile(r"^ShapeMismatch: expected (\[[^\]]*\]), actual (\[[^\]]*\]): (.*)$", re.S),
     "Shape mismatch in {2}: expected {0}, got {1}"),
    (re.compile(r"^NumericalInstability: (.*)$", re.S), "Numerical instability: {0}"),
)


def _display(err: FerroError) -> str:
    """The reference's Display text of a FerroError from its C-ABI message."""
    msg = str(err)
    for pat, fmt in _PATTERNS:
        m = pat.match(msg)
        if m:
            return fmt.format(*m.groups())
    return msg


def _call(fn, *args):
    try:
        return fn(*args)
    except FerroError as e:
        raise ValueError(_display(e)) from e


def _f64_2d(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _labels(y):
    # numpy1_to_ndarray_usize: i64 -> usize, which the labels round-trip
    return np.ascontiguousarray(np.asarray(y, dtype=np.int64))


class _Fitted:
    _fitted = None

    def _need(self):
        if self._fitted is None:
            raise RuntimeError("not fitted")
        return self._fitted


class _RsKMeans(_Fitted):
    def __init__(self, n_clusters=8, max_iter=300, tol=1e-4, n_init=10, random_state=None):
        self.n_clusters, self.max_iter, self.tol = int(n_clusters), int(max_iter), float(tol)
        self.n_init, self.random_state = int(n_init), random_state
        self._fitted = None

    def fit(self, x):
        m = (KMeans(self.n_clusters).with_max_iter(self.max_iter).with_tol(self.tol)
             .with_n_init(self.n_init).with_rng("stdrng"))
        if self.random_state is not None:
            m = m.with_random_state(int(self.random_state))
        self._fitted = _call(m.fit, _f64_2d(x))

    def predict(self, x):
        f = self._need()
        return np.asarray(_call(f.predict, _f64_2d(x)), dtype=np.int64)

    def transform(self, x):
        f = self._need()
        return np.asarray(_call(f.transform, _f64_2d(x)), dtype=np.float64)

    @property
    def cluster_centers_(self):
        return np.asarray(self._need().cluster_centers(), dtype=np.float64)

    @property
    def labels_(self):
        return np.asarray(self._need().labels(), dtype=np.int64)

    @property
    def inertia_(self):
        return float(self._need().inertia())

    @property
    def n_iter_(self):
        return int(self._need().n_iter())


class _RsMiniBatchKMeans(_Fitted):
    def __init__(self, n_clusters=8, max_iter=100, random_state=None):
        self.n_clusters, self.max_iter, self.random_state = int(n_clusters), int(max_iter), \
            random_state
        self._fitted = None

    def fit(self, x):
        m = MiniBatchKMeans(self.n_clusters).with_max_iter(self.max_iter).with_rng("stdrng")
        if self.random_state is not None:
            m = m.with_random_state(int(self.random_state))
        self._fitted = _call(m.fit, _f64_2d(x))

    def predict(self, x):
        f = self._need()
        return np.asarray(_call(f.predict, _f64_2d(x)), dtype=np.int64)

    @property
    def labels_(self):
        return np.asarray(self._need().labels(), dtype=np.int64)


class _RsKNeighborsClassifier(_Fitted):
    def __init__(self, n_neighbors=5):
        self.n_neighbors = int(n_neighbors)
        self._fitted = None

    def fit(self, x, y):
        m = KNeighborsClassifier().with_n_neighbors(self.n_neighbors)
        self._fitted = _call(m.fit, _f64_2d(x), _labels(y))

    def predict(self, x):
        f = self._need()
        return np.asarray(_call(f.predict, _f64_2d(x)), dtype=np.int64)

    @property
    def classes_(self):
        return np.asarray(self._need().classes(), dtype=np.int64)


class _RsKNeighborsRegressor(_Fitted):
    def __init__(self, n_neighbors=5):
        self.n_neighbors = int(n_neighbors)
        self._fitted = None

    def fit(self, x, y):
        m = KNeighborsRegressor().with_n_neighbors(self.n_neighbors)
        self._fitted = _call(m.fit, _f64_2d(x), np.asarray(y, dtype=np.float64))

    def predict(self, x):
        f = self._need()
        return np.asarray(_call(f.predict, _f64_2d(x)), dtype=np.float64)
