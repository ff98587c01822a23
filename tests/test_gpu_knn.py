"""Brute-force kNN parity on the GPU: the library against the oracle and the
golden vectors.  Neighbour indices and distances bit for bit in (distance,
index) order; votes and means bit for bit."""
import numpy as np
import pytest

import ferrolearn_amd as F
from oracle import oracle as O
from tests._golden import arr, bits, golden, ref_fixture

pytestmark = pytest.mark.gpu

G = golden()
PRECISIONS = ["fp64", "exact", "bf16x3", "bf16"]


def mixture(N, d, comps, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-5, 5, size=(comps, d))
    sd = rng.uniform(0.5, 2.0, size=comps)
    c = rng.integers(0, comps, size=N)
    X = means[c] + rng.normal(size=(N, d)) * sd[c, None]
    return X.astype(dtype), c


@pytest.mark.parametrize("precision", PRECISIONS)
def test_golden_knn(precision):
    for case in G["knn"]:
        data, Q, k = arr(case["data"]), arr(case["Q"]), case["k"]
        y, yr = np.array(case["y"]), arr(case["yr"])
        for w in (0, 1):
            W = F.Weights.Distance if w else F.Weights.Uniform
            clf = (F.KNeighborsClassifier().with_n_neighbors(k).with_weights(W)
                   .with_precision(precision).fit(data, y))
            dist, idx = clf.kneighbors(Q)
            for qi, res in enumerate(case["results"]):
                assert list(idx[qi]) == res["idx"]
                assert (bits(dist[qi]) == bits(arr(res["dist"]))).all()
            assert list(clf.predict(Q)) == [r[f"vote_w{w}"] for r in case["results"]]
            reg = (F.KNeighborsRegressor().with_n_neighbors(k).with_weights(W)
                   .with_precision(precision).fit(data, yr))
            want = arr([r[f"mean_w{w}"] for r in case["results"]])
            assert (bits(reg.predict(Q)) == bits(want)).all()


def test_reference_fixtures():
    fx = ref_fixture("kneighbors_classifier.json")
    X, y = np.array(fx["input"]["X"]), np.array(fx["input"]["y"])
    clf = F.KNeighborsClassifier().with_n_neighbors(5).fit(X, y)
    assert (clf.predict(X) == O.knn_classify(X, y, X, 5)).all()
    assert (clf.predict(X) == np.array(fx["expected"]["predictions"])).mean() >= 0.95
    fr = ref_fixture("kneighbors_regressor.json")
    X, y = np.array(fr["input"]["X"]), np.array(fr["input"]["y"])
    reg = F.KNeighborsRegressor().with_n_neighbors(5).fit(X, y)
    assert (bits(reg.predict(X)) == bits(O.knn_regress(X, y, X, 5))).all()
    fc = ref_fixture("sklearn_kneighbors_classifier_50x5.json")
    clf = F.KNeighborsClassifier().fit(np.array(fc["X_train"]), np.array(fc["y_train"]))
    assert (clf.predict(np.array(fc["X_test"])) == np.array(fc["predictions_test"])).all()


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("k", [1, 10, 33, 64])
def test_mixture_matches_oracle(precision, k):
    X, comp = mixture(20_000, 128, 64, seed=k)
    Q, _ = mixture(300, 128, 64, seed=1000 + k)
    y = comp % 10
    clf = (F.KNeighborsClassifier().with_n_neighbors(k).with_precision(precision)
           .fit(X, y))
    dist, idx = clf.kneighbors(Q)
    d_o, i_o = O.kneighbors(X.astype(np.float64), Q.astype(np.float64), k)
    assert (idx == i_o).all()
    assert (bits(dist) == bits(d_o)).all()
    assert (clf.predict(Q) == O.knn_classify(X.astype(np.float64), y,
                                             Q.astype(np.float64), k)).all()


@pytest.mark.parametrize("layout", ["interleaved", "contiguous"])
@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_duplicate_rows_exact_and_fallback(layout, precision):
    # Exact duplicates tie.  Spread over the database (interleaved) each split
    # keeps one copy per point and the certificate still holds; stored next
    # to each other (contiguous) one split fills its list with equal keys, the
    # certificate fails and the query takes the exact scan.  Either way the
    # answer is the (distance, index) order.
    rng = np.random.default_rng(5)
    base = rng.integers(-2, 3, size=(400, 4)).astype(np.float64)
    X = np.concatenate([base] * 30) if layout == "interleaved" else np.repeat(base, 30, axis=0)
    Q = rng.integers(-2, 3, size=(200, 4)).astype(np.float64)
    clf = (F.KNeighborsClassifier().with_n_neighbors(10).with_precision(precision)
           .fit(X, np.arange(len(X)) % 3))
    dist, idx = clf.kneighbors(Q)
    d_o, i_o = O.kneighbors(X, Q, 10)
    assert (idx == i_o).all() and (bits(dist) == bits(d_o)).all()
    if layout == "contiguous":
        assert clf.report["n_fallback"] > 0


def test_f64_queries_and_proba():
    X, comp = mixture(5000, 20, 8, seed=2, dtype=np.float64)
    Q, _ = mixture(100, 20, 8, seed=3, dtype=np.float64)
    y = comp * 7
    clf = (F.KNeighborsClassifier().with_n_neighbors(6).with_weights(F.Weights.Distance)
           .fit(X, y))
    pr = clf.predict_proba(Q)
    _, pr_o = O.knn_classify(X, y, Q, 6, weights=1, proba=True)
    assert (bits(pr) == bits(pr_o)).all()
    assert list(clf.classes()) == sorted(set(y.tolist()))
    assert clf.n_samples_fit() == 5000


def test_k_equals_n_and_errors():
    X = np.array([[1.0], [2.0], [3.0], [4.0], [5.0]])
    clf = F.KNeighborsClassifier().with_n_neighbors(5).fit(X, [0, 0, 0, 1, 1])
    assert list(clf.predict(X)) == [0] * 5
    with pytest.raises(F.InsufficientSamples):
        clf.kneighbors(X, 6)
    with pytest.raises(F.ShapeMismatch):
        clf.predict(np.ones((2, 3)))
    with pytest.raises(F.ShapeMismatch):
        clf.kneighbors(np.ones((2, 3)))


def test_device_arrays():
    X, comp = mixture(30_000, 96, 32, seed=8)
    Q, _ = mixture(500, 96, 32, seed=9)
    for prec in ("fp64", "bf16x3"):
        clf = F.KNeighborsClassifier().with_n_neighbors(10).with_precision(prec).fit(X, comp % 10)
        qd = F.to_device(Q)
        dist, idx = clf.kneighbors(qd)
        d_o, i_o = O.kneighbors(X.astype(np.float64), Q.astype(np.float64), 10)
        assert (idx.numpy() == i_o).all()
        assert (bits(dist.numpy()) == bits(d_o)).all()
        lab = clf.predict(qd).numpy()
        assert (lab == O.knn_classify(X.astype(np.float64), comp % 10,
                                      Q.astype(np.float64), 10)).all()


TORCH_INTEROP = r"""
import sys, numpy as np
import torch
torch.zeros(1, device="cuda")          # torch's HIP runtime first
sys.path.insert(0, sys.argv[1])
import ferrolearn_amd as F
from oracle import oracle as O
rng = np.random.default_rng(0)
X = rng.normal(size=(5000, 24)).astype(np.float32)
Q = rng.normal(size=(300, 24)).astype(np.float32)
clf = F.KNeighborsClassifier().with_n_neighbors(7).with_precision("bf16x3").fit(X, np.arange(5000) % 4)
dist, idx = clf.kneighbors(torch.from_numpy(Q).cuda())
d_o, i_o = O.kneighbors(X.astype(np.float64), Q.astype(np.float64), 7)
assert (idx.cpu().numpy() == i_o).all() and (dist.cpu().numpy() == d_o).all()
f = F.KMeans(5).with_n_init(1).with_max_iter(20).fit(torch.from_numpy(X).cuda(),
    init_centers=X[:5].astype(np.float64).reshape(1, 5, 24))
c_o, l_o, _, it_o = O.kmeans_lloyd(X.astype(np.float64), X[:5].astype(np.float64), 20, 1e-4)
assert (f.labels().cpu().numpy() == l_o).all() and f.n_iter() == it_o
# Strided views: columns 8..31 of wider tensors (the reads stay inside the
# storage), a step-2 row view and a transposed view must be packed first.
wideQ = torch.from_numpy(rng.normal(size=(300, 40)).astype(np.float32)).cuda()
wideQ[:, 8:32] = torch.from_numpy(Q).cuda()
dist, idx = clf.kneighbors(wideQ[:, 8:32])
assert (idx.cpu().numpy() == i_o).all() and (dist.cpu().numpy() == d_o).all()
dist, idx = clf.kneighbors(torch.from_numpy(np.ascontiguousarray(Q.T)).cuda().t())
assert (idx.cpu().numpy() == i_o).all()
Q2 = torch.from_numpy(np.repeat(Q, 2, axis=0)).cuda()[::2]
dist, idx = clf.kneighbors(Q2)
assert (idx.cpu().numpy() == i_o).all()
wideX = torch.zeros((5000, 33), dtype=torch.float32, device="cuda")
wideX[:, 1:25] = torch.from_numpy(X).cuda()
f = F.KMeans(5).with_n_init(1).with_max_iter(20).fit(wideX[:, 1:25],
    init_centers=X[:5].astype(np.float64).reshape(1, 5, 24))
assert (f.labels().cpu().numpy() == l_o).all() and f.n_iter() == it_o
router = F.CellRouter(clf, n_cells=12, cell_iter=5, hidden=64, seed=0)
rd, ri = router.kneighbors(wideQ[:, 8:32], n_probe=2, route="bounds")
assert (ri.cpu().numpy() == i_o).all() and (rd.cpu().numpy() == d_o).all()
try:
    clf.kneighbors(torch.zeros((2, 3, 24), device="cuda"))
    raise AssertionError("3-D input accepted")
except F.ShapeMismatch:
    pass
print("torch interop ok")
"""


def test_torch_tensors_in_fresh_process():
    # PyTorch ships its own HIP runtime; with torch initialised first the
    # library shares it and accepts CUDA tensors directly.
    import os
    import subprocess
    import sys
    pytest.importorskip("torch")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", TORCH_INTEROP, root], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torch interop ok" in r.stdout


@pytest.mark.parametrize("m,k,n", [(37, 19, 23), (1, 1, 1), (64, 16, 64), (65, 17, 129),
                                   (300, 0, 5), (200, 513, 170), (1000, 96, 1)])
def test_gemm_backend(m, k, n):
    # Backend::gemm = a.dot(b): f64 products, k summed in an unspecified
    # order, so each entry is within a few ulp of sum |a||b| of the exact one.
    import ctypes as C
    from ferrolearn_amd import _lib as L
    rng = np.random.default_rng(m * 7 + k * 3 + n)
    A = rng.normal(size=(m, k))
    B = rng.normal(size=(k, n))
    out = np.full((m, n), np.nan)
    st = L.lib().fl_backend_gemm_f64(A.ctypes.data_as(C.c_void_p), m, k,
                                     B.ctypes.data_as(C.c_void_p), k, n,
                                     out.ctypes.data_as(C.c_void_p))
    assert st == 0, L.last_error()
    ref = A @ B
    tol = (k + 2) * 2.0 ** -52 * (np.abs(A) @ np.abs(B)) + 1e-300
    assert (np.abs(out - ref) <= tol).all()


def test_sqrt_and_division_correctly_rounded():
    # The exact paths rely on IEEE sqrt and division on the GPU: distances of
    # random folds must equal the CPU's correctly rounded sqrt bit for bit.
    rng = np.random.default_rng(1)
    X = rng.normal(size=(4000, 3)) * np.exp(rng.uniform(-20, 20, size=(4000, 1)))
    Q = rng.normal(size=(50, 3))
    clf = F.KNeighborsClassifier().with_n_neighbors(64).with_precision("exact").fit(
        X, np.zeros(4000, np.int64))
    dist, idx = clf.kneighbors(Q)
    d_o, i_o = O.kneighbors(X, Q, 64)
    assert (idx == i_o).all() and (bits(dist) == bits(d_o)).all()


@pytest.mark.parametrize("precision", ["bf16x3", "fp64"])
def test_unsafe_magnitudes_fall_back_exactly(precision):
    # Values beyond the BF16 filter's range (and f64 overflow of the squares)
    # must still give the reference's answer: the library routes such data to
    # a filter that can take it.
    rng = np.random.default_rng(21)
    X = rng.normal(size=(3000, 16))
    X[7] *= 1e20
    X[11] *= 1e200                      # squares overflow to inf
    Q = rng.normal(size=(40, 16))
    Q[3] = X[7] * 1.0000001
    clf = (F.KNeighborsClassifier().with_n_neighbors(5).with_precision(precision)
           .fit(X, np.arange(3000) % 4))
    dist, idx = clf.kneighbors(Q)
    d_o, i_o = O.kneighbors(X, Q, 5)
    assert (idx == i_o).all() and (bits(dist) == bits(d_o)).all()


def test_nan_input_reports_numerical_instability():
    # The reference panics on NaN distances (kdtree.rs:386); the library
    # reports NumericalInstability instead of returning a wrong order.
    X = np.random.default_rng(0).normal(size=(500, 8))
    X[10, 3] = np.nan
    clf = F.KNeighborsClassifier().with_n_neighbors(3).fit(X, np.zeros(500, np.int64))
    with pytest.raises(F.NumericalInstability):
        clf.kneighbors(X[:20])
