"""KMeans parity on the GPU: the library against the oracle and the golden
vectors.  Labels, n_iter and centres must match bit for bit (centres follow
the reference's index-order sums); inertia within 1e-12 relative."""
import math

import numpy as np
import pytest

import ferrolearn_amd as F
from oracle import oracle as O
from tests._golden import arr, bits, golden, ref_fixture

pytestmark = pytest.mark.gpu

G = golden()


def blobs(n, d, k, seed, spread=1.0, box=10.0, dtype=np.float64):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-box, box, size=(k, d))
    X = cen[np.arange(n) % k] + rng.normal(scale=spread, size=(n, d))
    return X.astype(dtype)


def fit_lib(X, init, max_iter, tol, precision="fp64"):
    init = np.asarray(init, np.float64)
    k = init.shape[-2]
    m = (F.KMeans(k).with_n_init(1).with_max_iter(max_iter).with_tol(tol)
         .with_precision(precision))
    return m.fit(X, init_centers=init.reshape(1, k, -1))


@pytest.mark.parametrize("precision", ["fp64", "exact"])
def test_golden_lloyd_runs(precision):
    for run in G["lloyd"]:
        X, init = arr(run["X"]), arr(run["init"])
        f = fit_lib(X, init, run["max_iter"], run["tol"], precision)
        assert f.n_iter() == run["n_iter"]
        assert list(f.labels()) == run["labels"]
        assert (bits(f.cluster_centers()) == bits(arr(run["centers"]))).all()
        # serial-path sizes: the reference's row-order fold, bit for bit
        assert X.shape[0] * X.shape[1] < 100_000
        assert f.inertia().hex() == run["inertia"]


def test_golden_argmin_ties_predict():
    for case in G["argmin"]:
        C = arr(case["centers"])
        rows = arr(case["rows"])
        fitted = F.FittedKMeans(C, np.zeros(1, np.int64), 0.0, 0, {})
        assert list(fitted.predict(rows)) == case["labels"]


@pytest.mark.parametrize("k", [40, 100])
def test_filter_path_ties_lowest_index(k):
    # More than 32 centres takes the f64 MFMA filter + certificate path.
    rng = np.random.default_rng(k)
    C = rng.integers(-3, 4, size=(k, 6)).astype(np.float64)
    C[k - 1] = C[3]
    C[k - 2] = C[0]
    X = np.concatenate([rng.integers(-4, 5, size=(3000, 6)).astype(np.float64), C])
    fitted = F.FittedKMeans(C, np.zeros(1, np.int64), 0.0, 0, {})
    got = fitted.predict(X)
    want, _ = O.assign(X, C)
    assert (got == want).all()


@pytest.mark.parametrize("precision", ["bf16x3", "bf16", "fp8"])
def test_bf16_filter_ties_lowest_index(precision):
    # Integer data and duplicated initial centres: exact ties everywhere; the
    # argmin filter keeps the lowest index of equal keys per lane and the
    # re-score decides on exact folds.
    rng = np.random.default_rng(9)
    k = 70
    C = rng.integers(-3, 4, size=(k, 6)).astype(np.float64)
    C[k - 1] = C[3]
    C[k - 2] = C[0]
    C[40] = C[5]
    X = np.concatenate([rng.integers(-4, 5, size=(5000, 6)).astype(np.float64), C])
    c_o, l_o, i_o, it_o = O.kmeans_lloyd(X, C, 3, 0.0)
    f = fit_lib(X, C, 3, 0.0, precision)
    assert f.n_iter() == it_o
    assert (f.labels() == l_o).all()
    assert (bits(f.cluster_centers()) == bits(c_o)).all()


def test_sklearn_fixture_assign():
    fx = ref_fixture("kmeans.json")
    X = np.array(fx["input"]["X"])
    C = np.array(fx["expected"]["cluster_centers"])
    fitted = F.FittedKMeans(C, np.zeros(1, np.int64), 0.0, 0, {})
    assert (fitted.predict(X) == np.array(fx["expected"]["labels"])).all()
    f = fit_lib(X, C, 300, 1e-4)
    assert f.n_iter() == 1
    assert abs(f.inertia() - fx["expected"]["inertia"]) <= 1e-12 * fx["expected"]["inertia"]


@pytest.mark.parametrize("n,d,k,dtype,precision", [
    (100_000, 32, 16, np.float64, "fp64"),      # config-1 shape, direct exact path
    (40_000, 32, 64, np.float64, "fp64"),       # MFMA filter path
    (30_000, 64, 200, np.float32, "fp64"),      # f32 storage, MFMA filter path
    (20_000, 16, 48, np.float64, "exact"),
    (40_000, 32, 64, np.float64, "bf16x3"),     # BF16 MFMA argmin filter
    (30_000, 64, 200, np.float32, "bf16x3"),
    (30_000, 64, 200, np.float32, "bf16"),      # single product: more fallbacks
    (9_000, 128, 300, np.float32, "bf16x3"),
    (20_000, 20, 30, np.float32, "fp64"),       # f32 centroid sums, 8 lanes per row
    (20_000, 4, 10, np.float32, "bf16x3"),      # one lane per row
    (40_000, 32, 64, np.float64, "fp8"),        # FP8 E4M3 filter, BF16x3 for what it leaves
    (30_000, 64, 200, np.float32, "fp8"),
    (9_000, 128, 300, np.float32, "fp8"),
    (12_000, 130, 40, np.float32, "fp8"),       # d > 128: the FP8 form runs as BF16x3
    (3_000, 20, 12, np.float64, "fp64"),        # serial path (n*d < 100,000): row-order inertia
    (4_999, 20, 40, np.float32, "bf16x3"),
    (1_500, 64, 24, np.float32, "fp8"),
])
def test_lloyd_matches_oracle(n, d, k, dtype, precision):
    X = blobs(n, d, k, seed=n + k, spread=1.0, dtype=dtype)
    init = O.kmeanspp(X.astype(np.float64), k, 7)
    c_o, l_o, i_o, it_o = O.kmeans_lloyd(X.astype(np.float64), init, 100, 1e-4)
    f = fit_lib(X, init, 100, 1e-4, precision)
    assert f.n_iter() == it_o
    assert (f.labels() == l_o).all()
    assert (bits(f.cluster_centers()) == bits(c_o)).all()
    assert abs(f.inertia() - i_o) <= 1e-12 * i_o
    if n * d < 100_000:   # the reference's serial path folds the inertia in row order
        assert f.inertia() == i_o
    # inertia against the exact sum of the last assignment's folds, made
    # with the pre-update centres of the last iteration
    pre = O.kmeans_lloyd(X.astype(np.float64), init, it_o - 1, 1e-4)[0] if it_o > 1 else init
    _, dist = O.assign_rows(X.astype(np.float64), pre)
    assert abs(math.fsum(dist) - f.inertia()) <= 1e-12 * f.inertia()


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_overlapping_blobs_many_iterations(precision):
    # Overlapping clusters move for many iterations: one label flip anywhere
    # would change later centres and n_iter.
    X = blobs(20_000, 8, 40, seed=3, spread=6.0, box=5.0)
    init = O.kmeanspp(X, 40, 11)
    c_o, l_o, i_o, it_o = O.kmeans_lloyd(X, init, 60, 0.0)
    f = fit_lib(X, init, 60, 0.0, precision)
    assert it_o == 60 and f.n_iter() == 60
    assert (f.labels() == l_o).all()
    assert (bits(f.cluster_centers()) == bits(c_o)).all()


def test_n_init_keeps_strictly_smaller_inertia():
    X = blobs(5000, 4, 6, seed=9, spread=3.0)
    inits = np.stack([O.kmeanspp(X, 6, s) for s in (1, 2, 3)])
    c_o, l_o, i_o, it_o, best = O.kmeans_fit_from(X, inits, 100, 1e-4)
    f = F.KMeans(6).with_n_init(3).with_max_iter(100).fit(X, init_centers=inits)
    assert (f.labels() == l_o).all() and f.n_iter() == it_o
    assert (bits(f.cluster_centers()) == bits(c_o)).all()


def test_empty_cluster_keeps_centre_and_zero_iter():
    X = np.array([[0.0, 0.0], [0.1, 0.0], [10.0, 10.0], [10.0, 10.1]])
    init = np.array([[0.0, 0.0], [10.0, 10.0], [1e6, 1e6]])
    f = fit_lib(X, init, 10, 1e-4)
    c_o, l_o, i_o, it_o = O.kmeans_lloyd(X, init, 10, 1e-4)
    assert (bits(f.cluster_centers()) == bits(c_o)).all() and f.n_iter() == it_o
    f0 = fit_lib(X, init, 0, 1e-4)
    assert f0.n_iter() == 0 and (f0.labels() == 0).all()
    assert f0.inertia() == np.finfo(np.float64).max
    assert (f0.cluster_centers() == init).all()


def test_kmeanspp_matches_oracle():
    # many 4096-row blocks with a partial last one, and a large d that
    # takes the per-trial passes instead of the batched one
    for (n, d, k, seed) in ((3000, 5, 9, 42), (10_000, 16, 33, 7), (9000, 3, 1, 5),
                            (50_001, 8, 20, 3), (6000, 900, 4, 2)):
        X = blobs(n, d, max(k, 2), seed=seed)
        got = F.kmeans_plusplus(X, k, seed)
        want = O.kmeanspp(X, k, seed)
        assert (bits(got) == bits(want)).all(), (n, d, k)


def test_kmeanspp_zero_potential():
    # identical rows: the potential is 0 and the reference draws a random row
    X = np.ones((5000, 4))
    got = F.kmeans_plusplus(X, 5, 9)
    want = O.kmeanspp(X, 5, 9)
    assert (bits(got) == bits(want)).all()


def test_fit_without_init_uses_library_kmeanspp():
    X = blobs(8000, 5, 7, seed=1, spread=0.5)
    f = F.KMeans(7).with_n_init(2).with_random_state(5).with_max_iter(50).fit(X)
    inits = np.stack([O.kmeanspp(X, 7, 5), O.kmeanspp(X, 7, 6)])
    c_o, l_o, i_o, it_o, best = O.kmeans_fit_from(X, inits, 50, 1e-4)
    assert (f.labels() == l_o).all() and f.n_iter() == it_o
    assert (bits(f.cluster_centers()) == bits(c_o)).all()


def test_transform_and_predict_match_oracle():
    X = blobs(3000, 7, 5, seed=2)
    C = X[:50].copy()
    fitted = F.FittedKMeans(C, np.zeros(1, np.int64), 0.0, 0, {})
    assert (bits(fitted.transform(X)) == bits(O.transform(X, C))).all()
    assert (fitted.predict(X) == O.assign(X, C)[0]).all()
    assert (fitted.predict(X.astype(np.float32)) ==
            O.assign(X.astype(np.float32).astype(np.float64), C)[0]).all()


def test_device_array_fit():
    X = blobs(50_000, 32, 64, seed=4, dtype=np.float32)
    init = O.kmeanspp(X.astype(np.float64), 64, 3)
    c_o, l_o, i_o, it_o = O.kmeans_lloyd(X.astype(np.float64), init, 30, 1e-4)
    xd = F.to_device(X)
    f = (F.KMeans(64).with_n_init(1).with_max_iter(30)
         .fit(xd, init_centers=init.reshape(1, 64, 32)))
    assert f.n_iter() == it_o
    assert (f.labels().numpy() == l_o).all()
    assert (bits(f.cluster_centers().numpy()) == bits(c_o)).all()


def config3_blobs(n, d=64, comps=1024, seed=0):
    """The large-KMeans workload's data at reduced n: f32 rows from 1,024
    Gaussian blobs, means uniform in [-20, 20]^64, per-blob std uniform in
    [0.5, 3]."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(-20, 20, size=(comps, d))
    sd = rng.uniform(0.5, 3.0, size=comps)
    c = rng.integers(0, comps, size=n)
    return (means[c] + rng.normal(size=(n, d)) * sd[c, None]).astype(np.float32)


@pytest.fixture(scope="module")
def config3_case():
    X = config3_blobs(120_000)
    k = 1024
    init = F.kmeans_plusplus(X, k, 0)     # the library's k-means++, seed 0
    X64 = X.astype(np.float64)
    # the oracle's k-means++ picks the same rows (checked at this k on a
    # subsample, where the oracle finishes quickly)
    sub = X64[:6000]
    assert (bits(F.kmeans_plusplus(X[:6000], k, 0)) == bits(O.kmeanspp(sub, k, 0))).all()
    c_o, l_o, i_o, it_o = O.kmeans_lloyd(X64, init, 6, 0.0)
    return X, init, (c_o, l_o, i_o, it_o)


@pytest.mark.parametrize("precision", ["bf16x3", "fp64", "fp8"])
def test_config3_shape_k1024(config3_case, precision):
    # k = 1,024 on the large-KMeans workload's data: six Lloyd iterations with
    # tol 0 from the same initial centres; labels, n_iter and centres bit for
    # bit, inertia to 1e-12.
    X, init, (c_o, l_o, i_o, it_o) = config3_case
    f = fit_lib(F.to_device(X), init, 6, 0.0, precision)
    if precision == "fp8":
        print(f"fp8 at k=1024: {f.report['n_second']} of {6 * X.shape[0]} row-iterations "
              f"left to BF16x3, {f.report['n_fallback']} to the exact scan")
    assert f.n_iter() == it_o == 6
    assert (f.labels().numpy() == l_o).all()
    assert (bits(f.cluster_centers().numpy()) == bits(c_o)).all()
    assert abs(f.inertia() - i_o) <= 1e-12 * i_o


def test_kmeanspp_golden_vectors():
    # the library's k-means++ against the pure-Python restatement of
    # kmeans.rs:169-253 (same generator; n <= 4096 is the reference's own
    # left-fold order), including the zero-potential branch, k = 1 and k = n
    for case in G["kmeanspp"]:
        X = arr(case["X"])
        got = F.kmeans_plusplus(X, case["k"], int(case["seed"]))
        assert (bits(got) == bits(X[case["picks"]])).all(), (X.shape, case["k"])


def test_batched_kmeanspp_golden_vectors():
    # the batched path's device-side k-means++ (one restart, max_iter 0 keeps
    # the initial centres) against the same vectors
    items = []
    for case in G["kmeanspp"]:
        X = arr(case["X"])
        if X.shape[1] > 64:
            continue
        m = F.KMeans(case["k"]).with_n_init(1).with_max_iter(0).with_random_state(int(case["seed"]))
        items.append((F.to_device(X), m, None, X, case))
    fitted, _ = F.kmeans_fit_batched([(x, m, i) for x, m, i, _, _ in items])
    for f, (_, _, _, X, case) in zip(fitted, items):
        assert (bits(f.cluster_centers().numpy()) == bits(X[case["picks"]])).all()


@pytest.mark.parametrize("precision", ["fp64", "bf16x3", "exact"])
def test_many_clusters_beyond_lds_counts(precision):
    # more clusters than the update's LDS counters hold (8192): the member
    # order comes from a stable radix sort by label; every empty cluster
    # keeps its centre (kmeans.rs:387-392); no cluster limit but k <= n
    rng = np.random.default_rng(91)
    n, d, k = 30_000, 6, 9_000
    X = rng.normal(size=(n, d)) * 3.0
    init = X[rng.choice(n, k, replace=False)].copy()
    c_o, l_o, i_o, it_o = O.kmeans_lloyd(X, init, 4, 0.0)
    f = fit_lib(X, init, 4, 0.0, precision)
    assert f.n_iter() == it_o
    assert (f.labels() == l_o).all()
    assert (bits(f.cluster_centers()) == bits(c_o)).all()
    assert abs(f.inertia() - i_o) <= 1e-12 * i_o
    # predict with the fitted centres against the oracle's assignment
    lab, _ = O.assign_rows(X, f.cluster_centers())
    assert (f.predict(X) == lab).all()
