"""The ferrolearn-python `_Rs*` classes on the GPU (ferrolearn_amd._ferrolearn_rs):
results against the oracle, and an sklearn-style wrapper driving them the
way ferrolearn-python/python/ferrolearn's wrappers do (construct with the
estimator's parameters, fit on float64, read the fitted getters)."""
import pickle

import numpy as np
import pytest

import ferrolearn_amd as F
from ferrolearn_amd import _ferrolearn_rs as R
from oracle import oracle as O
from tests._golden import bits

from sklearn.base import BaseEstimator, ClusterMixin, TransformerMixin, clone
from sklearn.utils.validation import check_is_fitted, validate_data

pytestmark = pytest.mark.gpu


def blobs(n, d, k, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, size=(k, d))
    lab = rng.integers(0, k, size=n)
    return c[lab] + rng.normal(scale=spread, size=(n, d)), lab


def test_rs_kmeans_matches_oracle():
    X, _ = blobs(4000, 6, 5, 1)
    rs = R._RsKMeans(n_clusters=5, max_iter=60, n_init=2, random_state=3)
    rs.fit(X)
    c_o, l_o, i_o, it_o = O.kmeans_fit_stdrng(X, 5, 60, 1e-4, 2, 3)   # the reference's stream
    assert (rs.labels_ == l_o).all() and rs.labels_.dtype == np.int64
    assert rs.n_iter_ == it_o and isinstance(rs.n_iter_, int)
    assert (bits(rs.cluster_centers_) == bits(c_o)).all()
    Q, _ = blobs(500, 6, 5, 2)
    assert (rs.predict(Q) == O.assign(Q, c_o)[0]).all()
    assert (bits(rs.transform(Q)) == bits(O.transform(Q, c_o))).all()
    assert abs(rs.inertia_ - i_o) <= 1e-12 * i_o   # device inertia: 1024-row blocks


def test_rs_neighbors_and_minibatch_match_oracle():
    X, lab = blobs(3000, 8, 6, 5)
    Q, _ = blobs(400, 8, 6, 6)
    y = (lab % 3) * 7 - 4            # negative labels round-trip as in i64 -> usize -> i64
    clf = R._RsKNeighborsClassifier(n_neighbors=7)
    clf.fit(X, y)
    assert (clf.predict(Q) == O.knn_classify(X, y, Q, 7)).all()
    assert list(clf.classes_) == sorted(set(y.tolist()))
    yr = np.random.default_rng(1).normal(size=len(X))
    reg = R._RsKNeighborsRegressor()
    reg.fit(X, yr)
    assert (bits(reg.predict(Q)) == bits(O.knn_regress(X, yr, Q, 5))).all()
    mb = R._RsMiniBatchKMeans(n_clusters=4, max_iter=15, random_state=2)
    mb.fit(X)
    c_o, l_o, _, _ = O.minibatch_fit_stdrng(X, 4, 1024, 15, 0.0, 3, 2)
    assert (mb.labels_ == l_o).all()
    assert (mb.predict(Q) == O.assign(Q, c_o)[0]).all()


class KMeansWrapper(TransformerMixin, ClusterMixin, BaseEstimator):
    """An sklearn estimator over _RsKMeans in the shape of ferrolearn-python's
    KMeans wrapper (python/ferrolearn/_clusterers.py)."""

    def __init__(self, *, n_clusters=8, max_iter=300, tol=1e-4, n_init=10,
                 random_state=None):
        self.n_clusters, self.max_iter, self.tol = n_clusters, max_iter, tol
        self.n_init, self.random_state = n_init, random_state

    def _rs_new(self):
        return R._RsKMeans(n_clusters=self.n_clusters, max_iter=self.max_iter, tol=self.tol,
                           n_init=self.n_init, random_state=self.random_state)

    def fit(self, X, y=None):
        X = validate_data(self, X, dtype="float64")
        self._rs = self._rs_new()
        self._rs.fit(np.ascontiguousarray(X))
        self.cluster_centers_ = np.array(self._rs.cluster_centers_)
        self.labels_ = np.asarray(self._rs.labels_)
        self.inertia_ = float(self._rs.inertia_)
        self.n_iter_ = int(self._rs.n_iter_)
        self._fit_X = X.copy()
        return self

    def predict(self, X):
        check_is_fitted(self)
        return np.asarray(self._rs.predict(validate_data(self, X, reset=False,
                                                         dtype="float64")))

    def transform(self, X):
        check_is_fitted(self)
        return np.array(self._rs.transform(validate_data(self, X, reset=False,
                                                         dtype="float64")))

    def __getstate__(self):
        s = self.__dict__.copy()
        s.pop("_rs", None)
        return s

    def __setstate__(self, s):
        self.__dict__.update(s)
        self._rs = self._rs_new()
        self._rs.fit(self._fit_X)


def test_sklearn_style_wrapper():
    X, _ = blobs(2000, 4, 3, 9)
    est = KMeansWrapper(n_clusters=3, n_init=2, random_state=0)
    labels = clone(est).fit_predict(X)
    est.fit(X)
    assert (labels == est.labels_).all() and est.transform(X).shape == (2000, 3)
    back = pickle.loads(pickle.dumps(est))
    assert (back.predict(X) == est.predict(X)).all()
    with pytest.raises(ValueError, match="Invalid parameter `n_clusters`"):
        KMeansWrapper(n_clusters=0).fit(X)
