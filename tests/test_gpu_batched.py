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
