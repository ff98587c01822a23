"""Batched KMeans on the GPU: every problem of one fl_kmeans_fit_batched call
must equal the same fit run alone (labels, centres, inertia, n_iter bit for
bit), and a sample equals the oracle from the same initial centres."""
import numpy as np
import pytest

import ferrolearn_amd as F
from oracle import oracle as O
from tests._golden import bits

pytestmark = pytest.mark.gpu


def problem(seed):
    rng = np.random.default_rng(seed)
    n = int(np.exp(rng.uniform(np.log(500), np.log(20000))))
    d = int(rng.choice([8, 16, 32, 64]))
    k = int(np.exp(rng.uniform(np.log(4), np.log(64))))
    cen = rng.uniform(-10, 10, size=(k, d))
    X = cen[rng.integers(0, k, n)] + rng.normal(size=(n, d))
    dtype = np.float32 if seed % 2 else np.float64
    return X.astype(dtype), k


def test_batched_equals_one_by_one():
    probs, models, inits, xs = [], [], [], []
    for s in range(24):
        X, k = problem(s)
        prec = ["fp64", "bf16x3"][s % 2]
        m = (F.KMeans(k).with_n_init(2).with_max_iter(60).with_tol(1e-4)
             .with_random_state(100 + s).with_precision(prec))
        init = O.kmeanspp(X.astype(np.float64), k, s) if s % 3 == 0 else None
        if init is not None:
            init = np.concatenate([init, O.kmeanspp(X.astype(np.float64), k, s + 1)])
        xd = F.to_device(X)
        xs.append(X)
        probs.append((xd, m, init))
    fitted, rep = F.kmeans_fit_batched(probs, n_workers=6)
    assert rep["n_iter"] == sum(f.n_iter() for f in fitted)
    for (xd, m, init), f, X in zip(probs, fitted, xs):
        g = m.fit(xd, init_centers=init)
        assert f.n_iter() == g.n_iter()
        assert (f.labels().numpy() == g.labels().numpy()).all()
        assert (bits(f.cluster_centers().numpy()) == bits(g.cluster_centers().numpy())).all()
        assert f.inertia() == g.inertia()
    # the problems with explicit initial centres against the oracle
    for s in range(0, 24, 6):
        xd, m, init = probs[s]
        X = xs[s].astype(np.float64)
        c_o, l_o, i_o, it_o, best = O.kmeans_fit_from(X, init.reshape(2, m.n_clusters, -1),
                                                      60, 1e-4)
        f = fitted[s]
        assert f.n_iter() == it_o and (f.labels().numpy() == l_o).all()
        assert (bits(f.cluster_centers().numpy()) == bits(c_o)).all()


def test_batched_reports_the_failing_problem():
    X, k = problem(3)
    xd = F.to_device(X)
    probs = [(xd, F.KMeans(k), None), (xd, F.KMeans(0), None)]
    with pytest.raises(F.InvalidParameter, match="problem 1"):
        F.kmeans_fit_batched(probs)


def test_batched_edge_cases_equal_one_by_one():
    # max_iter 0 and 1, k = 1, k = n, identical rows (zero k-means++
    # potential), d off the register buckets (5, 24, 40), tol 0 with many
    # iterations, eight restarts, and d = 100 (outside the batched kernels:
    # a whole fit on a worker thread) in the same call.
    rng = np.random.default_rng(42)
    cases = []
    X = rng.normal(size=(3000, 5)); cases.append((X, F.KMeans(7).with_max_iter(0)))
    X = rng.normal(size=(2500, 8)).astype(np.float32); cases.append((X, F.KMeans(1).with_n_init(3)))
    X = rng.normal(size=(40, 3)); cases.append((X, F.KMeans(40).with_n_init(2)))
    X = np.ones((5000, 16)); cases.append((X, F.KMeans(4).with_n_init(2)))
    X = rng.normal(size=(9000, 24)); cases.append((X, F.KMeans(30).with_tol(0.0).with_max_iter(80)))
    X = (rng.normal(size=(20000, 40)) * 3).astype(np.float32)
    cases.append((X, F.KMeans(300).with_n_init(8).with_max_iter(25)))
    X = rng.normal(size=(3000, 100)); cases.append((X, F.KMeans(12).with_n_init(2)))
    X = rng.integers(-2, 3, size=(6000, 6)).astype(np.float64)   # exact ties everywhere
    cases.append((X, F.KMeans(50).with_n_init(4).with_max_iter(1)))
    probs = [(F.to_device(X), m.with_random_state(7 + i), None) for i, (X, m) in enumerate(cases)]
    fitted, rep = F.kmeans_fit_batched(probs)
    for i, ((xd, m, _), f) in enumerate(zip(probs, fitted)):
        g = m.fit(xd)
        assert f.n_iter() == g.n_iter(), i
        assert (f.labels().numpy() == g.labels().numpy()).all(), i
        assert (bits(f.cluster_centers().numpy()) == bits(g.cluster_centers().numpy())).all(), i
        assert bits(np.array([f.inertia()])) == bits(np.array([g.inertia()])), i
    assert fitted[0].n_iter() == 0 and fitted[0].inertia() == np.finfo(np.float64).max
    # a fit from the library's k-means++ (ten restarts, seeds 11..20) against
    # the oracle's k-means++ + Lloyd + strict-< choice of the restart
    X = cases[4][0]
    inits = np.stack([O.kmeanspp(X, 30, 11 + r) for r in range(10)])
    c_o, l_o, i_o, it_o, _ = O.kmeans_fit_from(X, inits, 80, 0.0)
    f = fitted[4]
    assert f.n_iter() == it_o and (f.labels().numpy() == l_o).all()
    assert (bits(f.cluster_centers().numpy()) == bits(c_o)).all()


def test_tuning_table_roundtrip_and_results_independent():
    from ferrolearn_amd import tuning
    t0 = tuning.get_raw_table()
    X, k = problem(5)
    xd = F.to_device(X)
    m = F.KMeans(k).with_n_init(3).with_max_iter(30).with_random_state(1)
    ref, _ = F.kmeans_fit_batched([(xd, m, None)])
    try:
        for v in tuning.variants_for(X.shape[1]):
            tuning.set_table({(X.shape[1], k): v})
            got, _ = F.kmeans_fit_batched([(xd, m, None)])
            assert (got[0].labels().numpy() == ref[0].labels().numpy()).all(), v
            assert (bits(got[0].cluster_centers().numpy())
                    == bits(ref[0].cluster_centers().numpy())).all(), v
    finally:
        tuning.set_raw_table(t0)
