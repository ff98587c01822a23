"""Unsupervised neighbours, the k-neighbours graph and large k on the GPU.

NearestNeighbors (ferrolearn-neighbors/src/nearest_neighbors.rs) and
kneighbors_graph (graph.rs) against the oracle, bit for bit; any
1 <= k <= n (knn.rs:288-301) through the exact sorted path; the few
queries a filter leaves; routed searches beyond the routed scan's k."""
import numpy as np
import pytest
import torch

import ferrolearn_amd as F
from oracle import oracle as O
from tests._golden import bits

pytestmark = pytest.mark.gpu


def mixture(N, d, comps, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-5, 5, size=(comps, d))
    sd = rng.uniform(0.5, 2.0, size=comps)
    c = rng.integers(0, comps, size=N)
    X = means[c] + rng.normal(size=(N, d)) * sd[c, None]
    return X.astype(dtype), c


def sample_data():
    """nearest_neighbors.rs:447-453."""
    return np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [5.0, 5.0], [6.0, 5.0], [5.0, 6.0]])


# ------------------------------------------------------------------ large k --

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_large_k_matches_oracle(dtype, precision):
    X, c = mixture(3000, 20, 12, 7, dtype)
    Q, _ = mixture(150, 20, 12, 8, dtype)
    y = c % 5
    yr = np.random.default_rng(3).normal(size=len(X))
    for k in (65, 200, len(X)):
        clf = (F.KNeighborsClassifier().with_n_neighbors(k).with_precision(precision)
               .fit(X, y))
        dist, idx = clf.kneighbors(Q)
        d_o, i_o = O.kneighbors(X, Q, k)
        assert (idx == i_o).all(), k
        assert (bits(dist) == bits(d_o)).all(), k
        assert (clf.predict(Q) == O.knn_classify(X, y, Q, k)).all(), k
        for w in (0, 1):
            W = F.Weights.Distance if w else F.Weights.Uniform
            reg = (F.KNeighborsRegressor().with_n_neighbors(k).with_weights(W)
                   .with_precision(precision).fit(X, yr))
            assert (bits(reg.predict(Q)) == bits(O.knn_regress(X, yr, Q, k, w))).all(), (k, w)


def test_large_k_device_queries_and_ties():
    # duplicated rows make equal distances: (distance, index) order
    rng = np.random.default_rng(11)
    base = rng.normal(size=(400, 6))
    X = np.concatenate([base, base, base[:100]])
    Q = np.concatenate([base[:30], rng.normal(size=(30, 6))])
    fitted = F.KNeighborsRegressor().with_n_neighbors(150).fit(X, np.zeros(len(X)))
    dist, idx = fitted.kneighbors(torch.from_numpy(Q).cuda())
    d_o, i_o = O.kneighbors(X, Q, 150)
    assert (idx.cpu().numpy() == i_o).all()
    assert (bits(dist.cpu().numpy()) == bits(d_o)).all()


def test_few_uncertified_queries_take_the_sorted_path():
    # exact duplicates of database rows at equal distance from a query:
    # the BF16 certificate cannot separate them, so those queries go to the
    # exact sorted fallback; every answer equals the oracle's
    rng = np.random.default_rng(2)
    X = rng.normal(size=(20000, 32)).astype(np.float32)
    X[5000:5010] = X[100]
    Q = np.concatenate([X[100:101] + 0.25, rng.normal(size=(300, 32)).astype(np.float32)])
    clf = F.KNeighborsClassifier().with_n_neighbors(5).with_precision("bf16x3").fit(
        X, np.arange(len(X)) % 3)
    dist, idx = clf.kneighbors(Q)
    assert clf.report["n_fallback"] >= 1
    d_o, i_o = O.kneighbors(X, Q, 5)
    assert (idx == i_o).all() and (bits(dist) == bits(d_o)).all()


def test_large_k_nan_reports_numerical_instability():
    X = np.random.default_rng(0).normal(size=(300, 4))
    X[7, 1] = np.nan
    fitted = F.KNeighborsClassifier().with_n_neighbors(100).fit(X, np.zeros(300, np.int64))
    with pytest.raises(F.NumericalInstability):
        fitted.kneighbors(X[:5])


# --------------------------------------------------------- NearestNeighbors --

def test_nearest_neighbors_reference_cases():
    x = sample_data()
    f = F.NearestNeighbors().fit(x)                     # test_fit_basic
    assert f.n_samples_fit() == 6 and f.shape() == (6, 2)
    f3 = F.NearestNeighbors().with_n_neighbors(3).fit(x)
    dist, idx = f3.kneighbors(x)                        # test_kneighbors_default_k
    assert dist.shape == (6, 3) and idx.shape == (6, 3)
    assert (np.diff(dist, axis=1) >= 0).all()
    assert f3.kneighbors(x, 2)[0].shape == (6, 2)       # test_kneighbors_override_k
    f1 = F.NearestNeighbors().with_n_neighbors(1).fit(x)
    d1, i1 = f1.kneighbors(x)                           # test_kneighbors_k1_self_match
    assert (i1[:, 0] == np.arange(6)).all() and (d1[:, 0] < 1e-10).all()
    with pytest.raises(F.ShapeMismatch):                # test_kneighbors_shape_mismatch
        f.kneighbors(np.array([[1.0, 2.0, 3.0]]))
    with pytest.raises(F.InvalidParameter, match="exceeds number of training samples"):
        f3.kneighbors(x, 100)                           # test_kneighbors_k_too_large
    with pytest.raises(F.InvalidParameter):             # test_kneighbors_k_zero
        f3.kneighbors(x, 0)
    for alg in (F.Algorithm.BruteForce, F.Algorithm.KdTree, F.Algorithm.BallTree):
        fa = F.NearestNeighbors().with_n_neighbors(3).with_algorithm(alg).fit(x)
        da, ia = fa.kneighbors(x)                       # test_all_algorithms_agree_kneighbors
        assert (ia == idx).all() and (bits(da) == bits(dist)).all()
    fs = F.NearestNeighbors().with_n_neighbors(1).fit(np.array([[3.0, 4.0]]))
    ds, is_ = fs.kneighbors(np.array([[0.0, 0.0]]))     # test_single_point
    assert is_[0, 0] == 0 and ds[0, 0] == 5.0
    f32 = F.NearestNeighbors().with_n_neighbors(2).fit(x.astype(np.float32))
    assert f32.kneighbors(x.astype(np.float32))[1].shape == (6, 2)   # test_f32_support


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_nearest_neighbors_matches_oracle(precision):
    X, _ = mixture(5000, 24, 16, 4)
    Q, _ = mixture(400, 24, 16, 5)
    f = F.NearestNeighbors().with_n_neighbors(10).with_precision(precision).fit(X)
    for k in (None, 1, 33, 100):
        dist, idx = f.kneighbors(Q, k)
        d_o, i_o = O.kneighbors(X, Q, 10 if k is None else k)
        assert (idx == i_o).all() and (bits(dist) == bits(d_o)).all(), k


# ------------------------------------------------------------------ graphs --

def check_graph(g, dist, idx, n_cols, distance_mode):
    indptr, indices, data, shape = O.knn_to_csr(dist, idx, n_cols, distance_mode)
    gi, gx, gd = (a.cpu().numpy() if torch.is_tensor(a) else a
                  for a in (g.indptr, g.indices, g.data))
    assert g.shape == shape
    assert (gi == indptr).all() and (gx == indices).all() and (bits(gd) == bits(data)).all()


@pytest.mark.parametrize("k", [3, 10, 70])
def test_kneighbors_graph_matches_oracle(k):
    X, c = mixture(2500, 16, 10, 9)
    Q, _ = mixture(300, 16, 10, 10)
    f = F.NearestNeighbors().with_n_neighbors(k).fit(X)
    d_o, i_o = O.kneighbors(X, Q, k)
    for mode in (F.GraphMode.Connectivity, F.GraphMode.Distance):
        check_graph(f.kneighbors_graph(Q, None, mode), d_o, i_o, len(X),
                    mode == F.GraphMode.Distance)
        gdev = f.kneighbors_graph(torch.from_numpy(Q).cuda(), k, mode)
        assert gdev.indices.is_cuda
        check_graph(gdev, d_o, i_o, len(X), mode == F.GraphMode.Distance)
    # the method forms of the supervised models (graph.rs:273-304)
    clf = F.KNeighborsClassifier().with_n_neighbors(k).fit(X, c % 3)
    check_graph(clf.kneighbors_graph(Q, None, F.GraphMode.Distance), d_o, i_o, len(X), True)
    reg = F.KNeighborsRegressor().with_n_neighbors(k).fit(X, np.zeros(len(X)))
    check_graph(reg.kneighbors_graph(Q, k, F.GraphMode.Connectivity), d_o, i_o, len(X), False)


def test_kneighbors_graph_free_function_and_api_proof():
    # ferrolearn-neighbors/tests/api_proof.rs:38-46, 73-80, 122-123
    x = np.array([[0.0, 0.0], [0.5, 0.0], [0.0, 0.5], [5.0, 5.0], [5.5, 5.0], [5.0, 5.5]])
    y = np.array([0, 0, 0, 1, 1, 1])
    f = F.KNeighborsClassifier().with_n_neighbors(3).with_weights(F.Weights.Distance).fit(x, y)
    g = f.kneighbors_graph(x, None, F.GraphMode.Connectivity)
    assert g.n_rows() == 6 and g.n_cols() == 6 and g.nnz() == 18
    assert f.kneighbors_graph(x, None, F.GraphMode.Distance).nnz() == 18
    r = F.KNeighborsRegressor().with_n_neighbors(3).fit(x, y.astype(float))
    assert r.kneighbors_graph(x, 2, F.GraphMode.Distance).nnz() == 12
    X, _ = mixture(1200, 8, 6, 12)
    g = F.kneighbors_graph(X, 7, F.GraphMode.Distance)  # graph.rs:47-57
    d_o, i_o = O.kneighbors(X, X, 7)
    check_graph(g, d_o, i_o, len(X), True)
    dense = g.to_dense()
    assert (np.diag(dense) == 0).all()                  # each row's own point, distance 0
    assert (g.to_scipy().toarray() == dense).all()
    empty = F.NearestNeighbors().with_n_neighbors(2).fit(X).kneighbors_graph(X[:0], 2)
    assert list(empty.indptr) == [0] and empty.nnz() == 0
    with pytest.raises(F.InvalidParameter):
        F.NearestNeighbors().with_n_neighbors(2).fit(X).kneighbors_graph(X, 0)


# ---------------------------------------------------------------- routed k --

def test_routed_k_beyond_scan_lists_is_the_brute_force():
    rng = np.random.default_rng(6)
    X = rng.normal(size=(8000, 10)).astype(np.float32)
    fitted = F.KNeighborsRegressor().with_n_neighbors(10).fit(X, np.zeros(len(X)))
    r = F.CellRouter(fitted, n_cells=8, cell_iter=4, hidden=64, seed=2)
    Q = rng.normal(size=(200, 10)).astype(np.float32)
    for k in (33, 90):
        rd, ri = r.kneighbors(Q, n_neighbors=k, n_probe=2, route="bounds")
        d_o, i_o = O.kneighbors(X, Q, k)
        assert (ri == i_o).all() and (bits(rd) == bits(d_o)).all(), k
        assert r.report["n_fallback"] == len(Q)
