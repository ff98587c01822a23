"""SpectralClustering with its K-Means step on the GPU.

Mirrors ferrolearn-cluster/src/spectral.rs: the builder (:68-118), the
validation of fit (:231-258) and its five steps (:262-297).  Steps 1-4 (RBF
affinity, normalised Laplacian, top eigenvectors, row normalisation) run on
the host in numpy in the reference's operation order; the eigenvectors come
from LAPACK rather than faer, so the embedding -- and with it the labels --
can differ from the reference's wherever an eigenvector's sign or a
degenerate eigenspace's basis does.  Step 5 is the reference's call
KMeans::new(k).with_n_init(n_init)[.with_random_state(seed)], fit, predict
(:283-291), made on this library's KMeans with the reference's StdRng
stream.
"""
from __future__ import annotations

import copy

import numpy as np

from .cluster import KMeans, _host2d, _is_device, _to_host
from .errors import InsufficientSamples, InvalidParameter


def _seq_sum(a, axis):
    """Left-to-right sums along an axis (Iterator::sum), not pairwise."""
    return np.cumsum(a, axis=axis).take(-1, axis=axis)


def affinity_matrix(x, gamma):
    """spectral.rs:148-163: 1 on the diagonal, else exp(-gamma * fold of
    (a - b)^2) with the fold in the input's type, widened to f64."""
    n, d = x.shape
    acc = np.zeros((n, n), x.dtype)
    for k in range(d):
        diff = x[:, None, k] - x[None, :, k]
        acc = acc + diff * diff
    a = np.exp(-float(gamma) * acc.astype(np.float64))
    np.fill_diagonal(a, 1.0)
    return a


def normalized_laplacian(a):
    """spectral.rs:166-180: D^-1/2 A D^-1/2 with zero degrees mapped to 0."""
    deg = _seq_sum(a, 1)
    dis = np.where(deg > 0.0, 1.0 / np.sqrt(np.where(deg > 0.0, deg, 1.0)), 0.0)
    return (dis[:, None] * a) * dis[None, :]


def top_k_eigenvectors(sym, k):
    """spectral.rs:183-200: eigenvectors of the k largest eigenvalues, in
    ascending eigenvalue order."""
    _, vecs = np.linalg.eigh(sym)
    return np.ascontiguousarray(vecs[:, vecs.shape[1] - k:])


def row_normalize(m):
    """spectral.rs:203-215: rows over their L2 norm; zero rows unchanged."""
    norm = np.sqrt(_seq_sum(m * m, 1))
    return np.where(norm[:, None] > 0.0, m / np.where(norm > 0.0, norm, 1.0)[:, None], m)


class FittedSpectralClustering:
    def __init__(self, labels, embedding):
        self._labels = labels
        self.embedding = embedding

    def labels(self):
        return self._labels


class SpectralClustering:
    """spectral.rs:68-118.  Defaults: gamma 1.0, n_init 10, random_state
    None.  `precision` selects the K-Means candidate filter (labels exact)."""

    def __init__(self, n_clusters: int):
        self.n_clusters = int(n_clusters)
        self.gamma = 1.0
        self.n_init = 10
        self.random_state = None
        self.precision = "fp64"

    def _with(self, **kw):
        o = copy.copy(self)
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def with_gamma(self, gamma: float):
        return self._with(gamma=float(gamma))

    def with_n_init(self, n_init: int):
        return self._with(n_init=int(n_init))

    def with_random_state(self, seed: int):
        return self._with(random_state=int(seed))

    def with_precision(self, precision: str):
        return self._with(precision=precision)

    def embed(self, x):
        """Steps 1-4 of fit (spectral.rs:265-281), in the input's type."""
        a = affinity_matrix(x, self.gamma)
        lap = normalized_laplacian(a)
        return row_normalize(top_k_eigenvectors(lap, self.n_clusters)).astype(x.dtype)

    def kmeans(self):
        """The K-Means of step 5 (spectral.rs:283-287)."""
        km = (KMeans(self.n_clusters).with_n_init(self.n_init).with_precision(self.precision)
              .with_rng("stdrng"))
        if self.random_state is not None:
            km = km.with_random_state(self.random_state)
        return km

    def fit(self, x, y=None) -> FittedSpectralClustering:
        x = _host2d(_to_host(x) if _is_device(x) else x)
        n = x.shape[0]
        if self.n_clusters == 0:
            raise InvalidParameter("InvalidParameter: n_clusters: must be at least 1")
        if not self.gamma > 0.0:
            raise InvalidParameter("InvalidParameter: gamma: must be positive")
        if n == 0 or n < self.n_clusters:
            raise InsufficientSamples(
                f"InsufficientSamples: required {self.n_clusters}, actual {n}: "
                "SpectralClustering requires at least n_clusters samples")
        e = self.embed(x)
        fitted = self.kmeans().fit(e)
        return FittedSpectralClustering(fitted.predict(e), e)

    def fit_predict(self, x):
        """spectral.rs:305-308."""
        return self.fit(x).labels()
