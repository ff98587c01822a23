"""Pins the CPU oracle (oracle/ref_cpu.c) before anything is checked against it.

Three independent sources:
  * golden vectors written by pure-Python binary64 loops (tests/golden/
    make_golden.py) -- bit-exact expectations;
  * the reference's own sklearn fixtures (fixtures/*.json and
    tests/fixtures/sklearn_reference/*.json of ferrolearn), with the
    tolerances the reference's oracle tests use;
  * closed forms and known answers from the reference's unit tests.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests._golden import arr, bits, golden, ref_fixture

G = golden()


# ----------------------------------------------------------------- fold -----
def test_fold_golden_bit_exact():
    for case in G["fold"]:
        a, b = arr(case["a"]), arr(case["b"])
        sq = O.sqeuclid(a, b)
        want = float.fromhex(case["sq"])
        assert bits([sq])[0] == bits([want])[0], (case, sq, want)
        assert bits([O.euclid(a, b)])[0] == bits([float.fromhex(case["dist"])])[0]


def test_fold_order_matters():
    # The two order-sensitive cases hold the same multiset of terms; a
    # restatement that reassociated would give them equal sums.
    c1, c2 = G["fold"][12], G["fold"][13]
    assert c1["sq"] != c2["sq"]


def test_fold_known_answers():
    # kdtree.rs:486-497: 3-4-5 triangle, self distance 0
    assert O.euclid([0.0, 0.0], [3.0, 4.0]) == 5.0
    assert O.euclid([1.0, 2.0, 3.0], [1.0, 2.0, 3.0]) == 0.0


def test_fold_is_not_fma():
    # (1+2^-30)^2 with an FMA keeps the 2^-60 term; two rounded ops lose it.
    a = [1.0 + 2.0 ** -30, 2.0 ** -60 * 0]
    b = [0.0, 0.0]
    t = 1.0 + 2.0 ** -30
    no_fma = t * t + 0.0
    assert O.sqeuclid(a, b) == no_fma


# --------------------------------------------------------------- argmin -----
def test_argmin_lowest_index_on_ties():
    for case in G["argmin"]:
        C = arr(case["centers"])
        for r, lab, dist in zip(case["rows"], case["labels"], case["dist"]):
            l, dv = O.nearest_center(arr(r), C)
            assert l == lab
            assert dv == float.fromhex(dist)
    # explicit: duplicated centre 0 and k-1 -> label 0
    C = np.array([[1.0, 1.0], [5.0, 5.0], [1.0, 1.0]])
    assert O.nearest_center([1.0, 1.0], C)[0] == 0


def test_argmin_nan_never_wins():
    # kmeans.rs:303: a NaN distance fails `<`, so the row keeps label 0 and
    # F::MAX when every distance is NaN.
    C = np.array([[np.nan, 0.0], [np.nan, 1.0]])
    lab, dist = O.nearest_center([0.0, 0.0], C)
    assert lab == 0 and dist == np.finfo(np.float64).max


# ---------------------------------------------------------------- Lloyd -----
def test_lloyd_golden_bit_exact():
    for run in G["lloyd"]:
        X, init = arr(run["X"]), arr(run["init"])
        C, L, inertia, it = O.kmeans_lloyd(X, init, run["max_iter"], run["tol"], nthreads=1)
        assert it == run["n_iter"]
        assert list(L) == run["labels"]
        assert (bits(C) == bits(arr(run["centers"]))).all()
        assert bits([inertia])[0] == bits([float.fromhex(run["inertia"])])[0]


def test_lloyd_zero_iterations_keeps_init():
    run = [r for r in G["lloyd"] if r["max_iter"] == 0][0]
    X, init = arr(run["X"]), arr(run["init"])
    C, L, inertia, it = O.kmeans_lloyd(X, init, 0, 1e-4)
    assert it == 0 and (L == 0).all() and inertia == np.finfo(np.float64).max
    assert (C == init).all()


def test_empty_cluster_keeps_old_centre():
    X = np.array([[0.0, 0.0], [0.1, 0.0], [10.0, 10.0]])
    old = np.array([[0.0, 0.0], [10.0, 10.0], [1e6, 1e6]])
    labels, _ = O.assign(X, old)
    new, counts, shift = O.recompute_centroids(X, labels, old)
    assert counts[2] == 0.0 and (new[2] == old[2]).all()
    assert new[0, 0] == 0.05


def test_centroid_sum_order_is_index_order():
    # Values chosen so that summing in a different order rounds differently.
    X = np.array([[1e16], [1.0], [1.0], [-1e16]])
    labels = np.zeros(4, np.int64)
    new, _, _ = O.recompute_centroids(X, labels, np.zeros((1, 1)))
    s = 0.0
    for v in (1e16, 1.0, 1.0, -1e16):
        s = s + v
    assert new[0, 0] == s / 4.0


def test_parallel_assign_labels_match_serial():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(5000, 40))
    C = rng.normal(size=(9, 40))
    l1, i1 = O.assign(X, C, nthreads=1)
    l8, i8 = O.assign(X, C, nthreads=8)
    assert (l1 == l8).all()
    assert abs(i1 - i8) <= 1e-12 * i1
    labels, dist = O.assign_rows(X, C)
    assert (labels == l1).all()
    assert abs(math.fsum(dist) - i1) <= 1e-12 * i1


def test_single_cluster_centre_is_mean():
    # kmeans.rs:800-817: centre of (1,1),(2,2),(3,3),(4,4) is (2.5, 2.5)
    X = np.array([[1.0, 1.0], [2.0, 2.0], [3.0, 3.0], [4.0, 4.0]])
    C, L, inertia, it = O.kmeans_lloyd(X, X[:1], 300, 1e-4)
    assert (L == 0).all() and C[0, 0] == 2.5 and C[0, 1] == 2.5


def test_k_equals_n_inertia_zero():
    # kmeans.rs:783-798
    X = np.array([[0.0, 0.0], [5.0, 5.0], [10.0, 10.0]])
    C, L, inertia, it = O.kmeans_lloyd(X, X.copy(), 300, 1e-4)
    assert inertia == 0.0 and len(set(L.tolist())) == 3


def test_identical_points_inertia_zero():
    X = np.ones((4, 2))
    C, L, inertia, it = O.kmeans_lloyd(X, X[:1], 300, 1e-4)
    assert inertia == 0.0


def test_transform_is_sqrt_fold():
    rng = np.random.default_rng(5)
    X = rng.normal(size=(20, 3))
    C = rng.normal(size=(4, 3))
    T = O.transform(X, C)
    for i in range(20):
        for c in range(4):
            assert T[i, c] == math.sqrt(O.sqeuclid(X[i], C[c]))


# ------------------------------------------------------ sklearn fixtures ----
def _relabel_agreement(a, b):
    """Pairwise co-membership agreement (label permutation invariant)."""
    a, b = np.asarray(a), np.asarray(b)
    same_a = a[:, None] == a[None, :]
    same_b = b[:, None] == b[None, :]
    return (same_a == same_b).mean()


@pytest.mark.parametrize("name,nested", [("kmeans.json", True), ("sklearn_kmeans_50x5.json", False)])
def test_kmeans_sklearn_fixture_fixed_point(name, nested):
    fx = ref_fixture(name)
    if nested:
        X = np.array(fx["input"]["X"])
        exp = fx["expected"]
        centers, labels, inertia = (np.array(exp["cluster_centers"]), np.array(exp["labels"]),
                                    exp["inertia"])
    else:
        X = np.array(fx["X_train"])
        centers, labels, inertia = (np.array(fx["cluster_centers"]), np.array(fx["labels"]),
                                    fx["inertia"])
    # Assignment against sklearn's centres reproduces sklearn's labels ...
    L, _ = O.assign(X, centers)
    assert (L == labels).all()
    # ... the inertia to 1e-12 relative (sum of exact folds) ...
    _, dist = O.assign_rows(X, centers)
    assert abs(math.fsum(dist) - inertia) <= 1e-12 * inertia
    # ... and sklearn's centres are the means of their members: a Lloyd run
    # started there stops after one iteration with the same labels.
    C, L2, inertia2, it = O.kmeans_lloyd(X, centers, 300, 1e-4)
    assert it == 1 and (L2 == labels).all()
    assert np.abs(C - centers).max() <= 1e-12 * np.abs(centers).max()


@pytest.mark.parametrize("name,nested", [("kmeans.json", True), ("sklearn_kmeans_50x5.json", False)])
def test_kmeans_sklearn_fixture_statistical(name, nested):
    # ferrolearn-cluster/tests/oracle_tests.rs:94-112: inertia ratio in
    # [0.95, 1.05] and pairwise label agreement >= 0.99 from k-means++ init.
    fx = ref_fixture(name)
    if nested:
        X = np.array(fx["input"]["X"])
        k = fx["params"]["n_clusters"]
        exp_l, exp_i = fx["expected"]["labels"], fx["expected"]["inertia"]
    else:
        X = np.array(fx["X_train"])
        k = fx["params"]["n_clusters"]
        exp_l, exp_i = fx["labels"], fx["inertia"]
    inits = np.stack([O.kmeanspp(X, k, 42 + r) for r in range(10)])
    C, L, inertia, it, best = O.kmeans_fit_from(X, inits, 300, 1e-4)
    assert 0.95 <= inertia / exp_i <= 1.05
    assert _relabel_agreement(L, exp_l) >= 0.99


def test_knn_classifier_iris_fixture():
    # ferrolearn-neighbors/tests/oracle_tests.rs:51-95: >= 95 % prediction
    # match and accuracy within 0.05 of sklearn's.
    fx = ref_fixture("kneighbors_classifier.json")
    X, y = np.array(fx["input"]["X"]), np.array(fx["input"]["y"])
    k = fx["params"]["n_neighbors"]
    pred = O.knn_classify(X, y, X, k)
    exp = np.array(fx["expected"]["predictions"])
    assert (pred == exp).mean() >= 0.95
    assert abs((pred == y).mean() - fx["expected"]["accuracy"]) <= 0.05


def test_knn_regressor_diabetes_fixture():
    # oracle_tests.rs:97-138: max |diff| < 1.0 and R^2 within 0.1.
    fx = ref_fixture("kneighbors_regressor.json")
    X, y = np.array(fx["input"]["X"]), np.array(fx["input"]["y"])
    k = fx["params"]["n_neighbors"]
    pred = O.knn_regress(X, y, X, k)
    exp = np.array(fx["expected"]["predictions"])
    assert np.abs(pred - exp).max() < 1.0
    r2 = 1 - ((y - pred) ** 2).sum() / ((y - y.mean()) ** 2).sum()
    assert abs(r2 - fx["expected"]["r2"]) <= 0.1


def test_knn_sklearn_equivalence_50x5():
    fc = ref_fixture("sklearn_kneighbors_classifier_50x5.json")
    X, y = np.array(fc["X_train"]), np.array(fc["y_train"])
    k = fc["params"]["n_neighbors"]
    assert (O.knn_classify(X, y, X, k) == np.array(fc["predictions"])).all()
    assert (O.knn_classify(X, y, np.array(fc["X_test"]), k) ==
            np.array(fc["predictions_test"])).all()
    fr = ref_fixture("sklearn_kneighbors_regressor_50x5.json")
    X, y = np.array(fr["X_train"]), np.array(fr["y_train"])
    k = fr["params"]["n_neighbors"]
    np.testing.assert_allclose(O.knn_regress(X, y, X, k), fr["predictions"], rtol=1e-12)
    np.testing.assert_allclose(O.knn_regress(X, y, np.array(fr["X_test"]), k),
                               fr["predictions_test"], rtol=1e-12)


# -------------------------------------------------------------- kNN golden --
def test_knn_golden_bit_exact():
    for case in G["knn"]:
        data, Q, k = arr(case["data"]), arr(case["Q"]), case["k"]
        y = np.array(case["y"])
        yr = arr(case["yr"])
        dist, idx = O.kneighbors(data, Q, k)
        for qi, res in enumerate(case["results"]):
            assert list(idx[qi]) == res["idx"]
            assert (bits(dist[qi]) == bits(arr(res["dist"]))).all()
        for w in (0, 1):
            pred = O.knn_classify(data, y, Q, k, weights=w)
            assert list(pred) == [r[f"vote_w{w}"] for r in case["results"]]
            pr = O.knn_regress(data, yr, Q, k, weights=w)
            assert (bits(pr) == bits(arr([r[f"mean_w{w}"] for r in case["results"]]))).all()


def test_knn_known_answers():
    # knn.rs:1050-1064 k=1 memorises; :1066-1080 k=n predicts the mode
    X = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    assert list(O.knn_classify(X, [0, 1, 2, 3], X, 1)) == [0, 1, 2, 3]
    X = np.array([[1.0], [2.0], [3.0], [4.0], [5.0]])
    assert list(O.knn_classify(X, [0, 0, 0, 1, 1], X, 5)) == [0] * 5
    # knn.rs:1262-1275 mean of 3 nearest = 10.0
    X = np.array([[0.0], [1.0], [2.0], [3.0], [100.0]])
    assert O.knn_regress(X, [0.0, 10.0, 20.0, 30.0, 1000.0], [[1.0]], 3)[0] == 10.0
    # knn.rs:1277-1296 inverse-distance weights
    got = O.knn_regress(np.array([[0.0], [10.0]]), [0.0, 100.0], [[1.0]], 2, weights=1)[0]
    assert got == (1.0 * 0.0 + (1.0 / 9.0) * 100.0) / (1.0 + 1.0 / 9.0)
    # knn.rs:1395-1411 exact match takes all the weight
    got = O.knn_regress(np.array([[0.0], [1.0], [2.0]]), [10.0, 20.0, 30.0], [[1.0]], 3, 1)[0]
    assert got == 20.0
    # kdtree.rs:445-453 brute force simple
    idx, _ = O.knn_query(np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]]),
                         [0.1, 0.1], 2)
    assert idx[0] == 0


def test_knn_k_clamped_to_n():
    data = np.arange(6.0).reshape(3, 2)
    idx, dist = O.knn_query(data, [0.0, 0.0], 10)
    assert list(idx) == [0, 1, 2]


def test_knn_matches_full_sort():
    rng = np.random.default_rng(11)
    data = rng.integers(-2, 3, size=(300, 3)).astype(np.float64)   # many ties
    Q = rng.integers(-2, 3, size=(20, 3)).astype(np.float64)
    dist, idx = O.kneighbors(data, Q, 17)
    for qi in range(20):
        full = sorted((math.sqrt(O.sqeuclid(data[i], Q[qi])), i) for i in range(300))[:17]
        assert [i for _, i in full] == list(idx[qi])


def test_proba_rows_sum_to_one():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(50, 3))
    y = rng.integers(0, 4, size=50)
    _, pr = O.knn_classify(X, y, X[:10], 5, proba=True)
    np.testing.assert_allclose(pr.sum(1), 1.0, rtol=1e-15)


# ----------------------------------------------------------- MiniBatch ------
def test_minibatch_update_golden():
    for case in G["minibatch"]:
        X, C = arr(case["X"]), arr(case["C"])
        C2, cnt, sh = O.minibatch_update(X, case["bidx"], case["blab"], C, case["counts"])
        assert (bits(C2) == bits(arr(case["C_out"]))).all()
        assert list(cnt) == case["counts_out"]
        assert sh == float.fromhex(case["shift"])


def test_minibatch_run_reasonable():
    fx = ref_fixture("mini_batch_kmeans.json")
    X = np.array(fx["input"]["X"])
    k = fx["params"]["n_clusters"]
    best = None
    for r in range(3):
        init = O.kmeanspp(X, k, 42 + r * 1_000_003)
        C, L, inertia, it = O.minibatch_run(X, init, 50, 100, 0.0, 42 + r * 1_000_003)
        assert it == 100                # tol = 0 -> every run does max_iter
        if best is None or inertia < best:
            best = inertia
    # ferrolearn-cluster/tests/oracle_tests.rs:199-250: ratio in [0.9, 1.1]
    assert 0.9 <= best / fx["expected"]["inertia"] <= 1.1


# ----------------------------------------------------------- generator ------
def test_rng_known_answer_and_stream():
    # SplitMix64 seeded with 0 gives 0xe220a8397b1dcdaf first (published
    # known answer); the xoshiro256** stream must match the Python restatement.
    from tests.golden.make_golden import rng_stream
    x = 0x9E3779B97F4A7C15
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
    assert z ^ (z >> 31) == 0xE220A8397B1DCDAF
    for seed, vals in G["rng"].items():
        got = O.rng_stream(int(seed), 8)
        assert [int(v) for v in got] == [int(v) for v in vals]
        assert rng_stream(int(seed), 8) == [int(v) for v in vals]


def test_blocked_sum_and_search():
    rng = np.random.default_rng(1)
    v = rng.random(10000)
    s = O.blocked_sum(v)
    assert abs(s - math.fsum(v)) <= 1e-12 * s
    # two-level definition, restated
    ref = 0.0
    for b in range(0, 10000, 4096):
        p = 0.0
        for x in v[b:b + 4096]:
            p = p + x
        ref = ref + p
    assert s == ref
    thr = 0.37 * s
    i = O.blocked_search(v, thr)
    assert v[:i].sum() < thr * (1 + 1e-12) and v[:i + 1].sum() >= thr * (1 - 1e-12)
    assert O.blocked_search(v, 2 * s) == 9999


def test_kmeanspp_deterministic_and_picks_rows():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(500, 4))
    a = O.kmeanspp(X, 7, 123)
    b = O.kmeanspp(X, 7, 123)
    assert (a == b).all()
    rows = {tuple(r) for r in X}
    assert all(tuple(r) in rows for r in a)
    assert not (O.kmeanspp(X, 7, 124) == a).all()


def test_kmeanspp_spreads_over_blobs():
    # D^2 sampling on well separated blobs picks one centre per blob.
    rng = np.random.default_rng(4)
    cen = rng.uniform(-100, 100, size=(6, 3))
    X = np.concatenate([c + rng.normal(scale=0.1, size=(100, 3)) for c in cen])
    C = O.kmeanspp(X, 6, 7)
    blob = [int(np.argmin(((cen - c) ** 2).sum(1))) for c in C]
    assert sorted(blob) == list(range(6))


# ------------------------------------------------------------ router net -----
# The numpy restatement of the router network (mlp.rs) is pinned by central
# finite differences of its own loss (a dropped term, a transposed operand or
# a wrong mask in the analytic gradient fails these), and by the closed form
# of a zero network.
def _mlp_case(seed, d=5, h=6, c=7, B=9, k=3):
    rng = np.random.default_rng(seed)
    n = d * h + h + h * h + h + h * c + c
    p = rng.normal(scale=0.5, size=n)
    X = rng.normal(size=(B, d))
    cells = rng.integers(0, c, size=(B, k))
    return p, X, cells, d, h, c


@pytest.mark.parametrize("alpha", [0.0, 0.01])
def test_mlp_oracle_finite_differences(alpha):
    p, X, cells, d, h, c = _mlp_case(3)
    loss, g = O.mlp_loss_grad(p, X, cells, d, h, c, alpha=alpha)
    eps = 1e-6
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = eps
        lp, _ = O.mlp_loss_grad(p + e, X, cells, d, h, c)
        lm, _ = O.mlp_loss_grad(p - e, X, cells, d, h, c)
        # mlp.rs adds alpha W to grad_w without a matching loss term
        w_entries = np.concatenate([np.ones(d * h), np.zeros(h), np.ones(h * h), np.zeros(h),
                                    np.ones(h * c), np.zeros(c)])
        fd = (lp - lm) / (2 * eps) + alpha * p[i] * w_entries[i]
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_oracle_zero_network():
    d, h, c, B, k = 4, 3, 5, 6, 2
    rng = np.random.default_rng(0)
    n = d * h + h + h * h + h + h * c + c
    X = rng.normal(size=(B, d))
    cells = rng.integers(0, c, size=(B, k))
    loss, g = O.mlp_loss_grad(np.zeros(n), X, cells, d, h, c)
    # uniform softmax: loss ln c; grad_b3 = 1/c - mean target
    assert abs(loss - math.log(c)) < 1e-15
    y = np.zeros((B, c))
    for i in range(B):
        for cc in cells[i]:
            y[i, cc] += 1.0 / k
    gb3 = g[-c:]
    assert np.allclose(gb3, 1.0 / c - y.mean(axis=0), atol=1e-15)
    assert np.all(g[:-c] == 0.0)


def test_bf16_round_ties_to_even():
    v = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.0 - 2**-7 - 2**-9, 0.0])
    assert list(O.bf16_round(v)) == [1.0, 1.0, 1.0 + 2**-6, -2.0 - 2**-6, 0.0]


# ------------------------------------------- k-means++ / MiniBatch fits ------
# The golden generator restates kmeans.rs:169-253 and mini_batch_kmeans.rs:
# 410-592 in pure Python with the library's generator; these pin the C
# oracle's or_kmeanspp / or_minibatch_fit (first pick, n_trials = 2 +
# floor(ln k), D^2 thresholds, strict-< potential, the zero-potential branch,
# k = 1, batch >= n, run seeds base + r*1000003, row-order inertia) bit for
# bit.  A wrong n_trials or keeping the largest potential changes the picks.
def test_kmeanspp_golden_bit_exact():
    for case in G["kmeanspp"]:
        X = arr(case["X"])
        got = O.kmeanspp(X, case["k"], int(case["seed"]))
        want = X[case["picks"]]
        assert (bits(got) == bits(want)).all(), (X.shape, case["k"])


def test_kmeanspp_golden_is_sensitive():
    # the restatement itself: one trial instead of 2 + floor(ln k), or the
    # largest potential instead of the smallest, picks other rows
    from tests.golden import make_golden as MG
    case = G["kmeanspp"][1]
    X = [[float.fromhex(v) for v in r] for r in case["X"]]
    picks = MG.kmeanspp(X, case["k"], MG.Gen(int(case["seed"])))
    assert picks == case["picks"]
    orig_log = MG.math.log
    try:
        MG.math.log = lambda v: -10.0          # n_trials -> max(2 - 10, 1) = 1
        assert MG.kmeanspp(X, case["k"], MG.Gen(int(case["seed"]))) != case["picks"]
    finally:
        MG.math.log = orig_log


def test_minibatch_fit_golden_bit_exact():
    for case in G["minibatch_fit"]:
        X = arr(case["X"])
        c, l, inertia, it = O.minibatch_fit(X, case["k"], case["batch"], case["max_iter"],
                                            case["tol"], case["n_init"], case["seed"])
        assert it == case["n_iter"]
        assert list(l) == case["labels"]
        assert (bits(c) == bits(arr(case["centers"]))).all()
        assert inertia == float.fromhex(case["inertia"])


def test_knn_to_csr_pinned_by_its_dense_matrix():
    # graph.rs:91-123: the CSR with ascending columns of the matrix whose
    # row i holds value v at column idx[i, j] -- a CSR with sorted columns is
    # unique for a given dense matrix, so these three facts pin it
    rng = np.random.default_rng(0)
    nq, n, k = 40, 25, 6
    idx = np.stack([rng.permutation(n)[:k] for _ in range(nq)]).astype(np.int64)
    dist = rng.uniform(0.5, 3.0, size=(nq, k))
    for distance_mode in (False, True):
        indptr, indices, data, shape = O.knn_to_csr(dist, idx, n, distance_mode)
        assert shape == (nq, n) and (indptr == np.arange(nq + 1) * k).all()
        rows = indices.reshape(nq, k)
        assert (np.diff(rows, axis=1) > 0).all()
        dense = np.zeros(shape)
        dense[np.repeat(np.arange(nq), k), indices] = data
        want = np.zeros(shape)
        want[np.repeat(np.arange(nq), k), idx.ravel()] = dist.ravel() if distance_mode else 1.0
        assert (dense == want).all()


def test_pairwise_argmin_known_answers_and_integers():
    # pairwise.rs:732-738 (euclidean_basic): distances 4 and 5
    idx, mins = O.pairwise_argmin_min(np.array([[0.0, 0.0], [3.0, 0.0]]), np.array([[0.0, 4.0]]))
    assert list(idx) == [0, 0] and list(mins) == [4.0, 5.0]
    # small integers: every operation of the expansion is exact, so the
    # answer is the integer brute force with the lowest index among ties and
    # the correctly rounded root of the exact squared distance
    rng = np.random.default_rng(1)
    X = rng.integers(-6, 7, size=(60, 5)).astype(np.float64)
    Y = rng.integers(-6, 7, size=(23, 5)).astype(np.float64)
    Y[11] = Y[4]
    idx, mins = O.pairwise_argmin_min(X, Y)
    for i in range(len(X)):
        sq = [int(((X[i] - Y[j]) ** 2).sum()) for j in range(len(Y))]
        j = sq.index(min(sq))
        assert idx[i] == j and mins[i] == math.sqrt(sq[j])


def test_pairwise_argmin_cancellation_and_nan():
    # far from the origin the expansion cancels: a plain-Python restatement
    # of pairwise.rs:170-185 / :564-588 (floats are IEEE doubles) gives the
    # same bits, and differs from the squared_euclidean fold's answer
    rng = np.random.default_rng(2)
    X = 1e8 + rng.normal(size=(40, 3))
    Y = 1e8 + rng.normal(size=(30, 3))
    idx, mins = O.pairwise_argmin_min(X, Y)
    differs = False
    for i in range(len(X)):
        xs = 0.0
        for v in X[i]:
            xs = xs + float(v) * float(v)
        best, bj = 0.0, 0
        for j in range(len(Y)):
            ys = dot = 0.0
            for a, b in zip(X[i], Y[j]):
                ys = ys + float(b) * float(b)
                dot = dot + float(a) * float(b)
            v = math.sqrt(max((xs + ys) - 2.0 * dot, 0.0))
            if j == 0 or v < best:
                best, bj = v, j
        assert idx[i] == bj and mins[i] == best
        differs |= bj != int(np.argmin(((X[i] - Y) ** 2).sum(1)))
    assert differs
    # f64::max returns the non-NaN operand: a NaN row has distance 0 to all
    Xn = np.array([[np.nan, 1.0], [0.0, 0.0]])
    idx, mins = O.pairwise_argmin_min(Xn, np.array([[5.0, 5.0], [0.0, 1.0]]))
    assert list(idx) == [0, 1] and mins[0] == 0.0 and mins[1] == 1.0


# ---------------------------------------------------------------- StdRng ---

def _block_hex(fn, rounds, ctr, key=(0,) * 8, tail=(0, 0, 0)):
    st = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574, *key, ctr, *tail]
    return np.asarray(fn(st, rounds), dtype="<u4").tobytes().hex()


def test_chacha_block_published_vectors():
    # RFC 8439 section 2.3.2 (key 00..1f, counter 1, nonce 0:4a:0:0:0:9 words)
    # and the zero-key zero-nonce keystreams of ChaCha20 (blocks 0 and 1) and
    # ChaCha12 (block 0), both restatements of the block function
    from tests.golden import make_golden as MG
    key = [0x03020100, 0x07060504, 0x0B0A0908, 0x0F0E0D0C,
           0x13121110, 0x17161514, 0x1B1A1918, 0x1F1E1D1C]
    rfc = ("10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
           "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")
    zero20 = ("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
              "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    zero20b = ("9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
               "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f")
    zero12 = ("9bf49a6a0755f953811fce125f2683d50429c3bb49e074147e0089a52eae155f"
              "0564f879d27ae3c02ce82834acfa8c793a629f2ca0de6919610be82f411326be")
    for fn in (O.chacha_block, MG.StdRng.chacha):
        assert _block_hex(fn, 20, 1, key, (0x09000000, 0x4A000000, 0)) == rfc
        assert _block_hex(fn, 20, 0) == zero20
        assert _block_hex(fn, 20, 1) == zero20b
        assert _block_hex(fn, 12, 0) == zero12


def test_stdrng_draws_golden():
    for seed, words in G["stdrng_stream"].items():
        assert list(O.stdrng_stream(int(seed), len(words))) == words, seed
    r = G["stdrng_ranges"]
    got = O.stdrng_ranges(r["seed"], [int(v) for v in r["lo"]], [int(v) for v in r["hi"]])
    assert [str(v) for v in got] == r["out"]
    f = G["stdrng_f64"]
    assert [float.hex(float(v)) for v in O.stdrng_f64s(f["seed"], len(f["out"]))] == f["out"]


def test_stdrng_range_is_uniform_and_in_bounds():
    # Canon's method over a span that is not a power of two: every value in
    # [lo, hi) and none outside; counts near uniform
    lo, hi, m = 3, 17, 14000
    got = O.stdrng_ranges(5, [lo] * m, [hi] * m)
    assert got.min() == lo and got.max() == hi - 1
    counts = np.bincount(got - lo, minlength=hi - lo)
    assert counts.min() > 0.8 * m / (hi - lo) and counts.max() < 1.2 * m / (hi - lo)


def test_kmeanspp_stdrng_golden():
    for case in G["kmeanspp_stdrng"]:
        X = arr(case["X"])
        got = O.kmeanspp_stdrng(X, case["k"], int(case["seed"]))
        assert (bits(got) == bits(X[case["picks"]])).all(), (X.shape, case["k"])


def test_kmeans_and_minibatch_fit_stdrng_golden():
    for case in G["kmeans_fit_stdrng"]:
        X = arr(case["X"])
        c, l, inertia, it = O.kmeans_fit_stdrng(X, case["k"], case["max_iter"], case["tol"],
                                                case["n_init"], case["seed"])
        assert it == case["n_iter"] and list(l) == case["labels"]
        assert (bits(c) == bits(arr(case["centers"]))).all()
        assert inertia == float.fromhex(case["inertia"])
    for case in G["minibatch_fit_stdrng"]:
        X = arr(case["X"])
        c, l, inertia, it = O.minibatch_fit_stdrng(X, case["k"], case["batch"],
                                                   case["max_iter"], case["tol"],
                                                   case["n_init"], case["seed"])
        assert it == case["n_iter"] and list(l) == case["labels"]
        assert (bits(c) == bits(arr(case["centers"]))).all()
        assert inertia == float.fromhex(case["inertia"])


def _euclid_loops(X, Y):
    """pairwise.rs:166-183 as plain Python loops of binary64 operations."""
    xs = [0.0] * len(X)
    for i, x in enumerate(X):
        acc = 0.0
        for v in x:
            acc = acc + v * v
        xs[i] = acc
    ys = [0.0] * len(Y)
    for j, y in enumerate(Y):
        acc = 0.0
        for v in y:
            acc = acc + v * v
        ys[j] = acc
    out = []
    for i, x in enumerate(X):
        row = []
        for j, y in enumerate(Y):
            dot = 0.0
            for a, b in zip(x, y):
                dot = dot + a * b
            v = (xs[i] + ys[j]) - 2.0 * dot
            v = 0.0 if v != v else max(v, 0.0)   # f64::max: the non-NaN operand
            row.append(math.sqrt(v))
        out.append(row)
    return np.array(out)


def test_euclidean_distances_doc_example():
    # pairwise.rs:146-154
    d = O.euclidean_distances(np.array([[0.0, 0.0], [1.0, 0.0]]),
                              np.array([[0.0, 0.0], [0.0, 1.0]]))
    assert d[0, 0] == 0.0 and d[0, 1] == 1.0 and d[1, 0] == 1.0 and d[1, 1] == math.sqrt(2.0)


def test_euclidean_distances_against_loops():
    rng = np.random.default_rng(11)
    for n, m, d, scale in ((7, 5, 3, 1.0), (12, 9, 17, 1e8), (4, 6, 40, 1e-3)):
        X = rng.normal(size=(n, d)) * scale + scale
        Y = rng.normal(size=(m, d)) * scale + scale
        Y[0] = X[0]                     # a zero distance that the expansion can miss
        got = O.euclidean_distances(X, Y)
        assert (bits(got) == bits(_euclid_loops(X.tolist(), Y.tolist()))).all()
        # max(., 0) clamps the cancelled negatives, and the row minima are
        # pairwise_distances_argmin_min's
        assert (got >= 0.0).all()
        idx, mins = O.pairwise_argmin_min(X, Y)
        assert (bits(mins) == bits(got.min(axis=1))).all()
        assert (idx == got.argmin(axis=1)).all()


def test_euclidean_distances_nan_is_clamped_like_f64_max():
    # f64::max returns the non-NaN operand: a NaN expansion becomes 0
    X = np.array([[np.nan, 1.0], [1.0, 2.0]])
    Y = np.array([[0.0, 0.0]])
    d = O.euclidean_distances(X, Y)
    assert d[0, 0] == 0.0 and d[1, 0] == math.sqrt(5.0)


def test_bisecting_golden_bit_exact():
    # pure-Python restatement of bisecting_kmeans.rs:174-505 with its own
    # StdRng (tests/golden/make_golden.py)
    for run in golden()["bisecting"]:
        X = arr(run["X"])
        C, L, inertia = O.bisecting_kmeans(X, run["n_clusters"], run["max_iter"], run["n_init"],
                                           int(run["seed"]), run["strategy"])
        assert list(L) == run["labels"]
        assert (bits(C) == bits(arr(run["centers"]))).all()
        assert inertia.hex() == run["inertia"]


def test_bisecting_reference_unit_tests():
    # bisecting_kmeans.rs:612-700: three tight blobs split into the blobs,
    # two clusters, one cluster
    x = np.array([[0.0, 0.0], [0.1, 0.1], [-0.1, 0.1], [10.0, 10.0], [10.1, 10.1], [9.9, 10.1],
                  [0.0, 10.0], [0.1, 10.1], [-0.1, 9.9]])
    _, L, _ = O.bisecting_kmeans(x, 3, seed=42)
    assert L[0] == L[1] == L[2] and L[3] == L[4] == L[5] and L[6] == L[7] == L[8]
    assert len({L[0], L[3], L[6]}) == 3
    x2 = np.array([[0.0, 0.0], [0.5, 0.0], [0.0, 0.5], [0.5, 0.5], [10.0, 10.0], [10.5, 10.0],
                   [10.0, 10.5], [10.5, 10.5]])
    _, L2, _ = O.bisecting_kmeans(x2, 2, seed=42)
    assert len(set(L2[:4])) == 1 and len(set(L2[4:])) == 1 and L2[0] != L2[4]
    C1, L1, i1 = O.bisecting_kmeans(np.array([[1.0, 1.0], [2.0, 2.0], [3.0, 3.0], [4.0, 4.0]]), 1,
                                    seed=42)
    assert (L1 == 0).all() and C1.shape == (1, 2) and C1[0, 0] == 2.5 and i1 == 10.0


def test_bisecting_identical_rows_do_not_loop():
    with pytest.raises(RuntimeError, match="empty"):
        O.bisecting_kmeans(np.ones((10, 2)), 2, max_stall=50)
