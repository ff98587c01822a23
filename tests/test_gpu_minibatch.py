"""MiniBatchKMeans parity on the GPU against the oracle's restatement of
mini_batch_kmeans.rs (same generator, same batches): centres, labels, n_iter
and the row-order inertia fold must match bit for bit."""
import numpy as np
import pytest

import ferrolearn_amd as F
from oracle import oracle as O
from tests._golden import bits

pytestmark = pytest.mark.gpu


def blobs(n, d, k, seed, spread=1.0, box=10.0, dtype=np.float64):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-box, box, size=(k, d))
    X = cen[np.arange(n) % k] + rng.normal(scale=spread, size=(n, d))
    return X.astype(dtype)


def fit_mb(X, init, batch, max_iter, tol, seed, precision="fp64", n_init=1):
    init = np.asarray(init, np.float64)
    k = init.shape[-2]
    m = (F.MiniBatchKMeans(k).with_batch_size(batch).with_max_iter(max_iter).with_tol(tol)
         .with_n_init(n_init).with_random_state(seed).with_precision(precision))
    return m.fit(X, init_centers=init.reshape(n_init, k, -1))


@pytest.mark.parametrize("n,d,k,batch,max_iter,tol,dtype,precision", [
    (20_000, 8, 16, 1024, 50, 0.0, np.float64, "fp64"),     # defaults' shape, direct path
    (20_000, 8, 16, 256, 200, 1e-3, np.float64, "fp64"),    # early stop on the shift
    (500, 4, 6, 1024, 20, 0.0, np.float64, "fp64"),         # batch >= n: full shuffles
    (30_000, 32, 48, 1024, 40, 0.0, np.float32, "fp64"),    # f32, filter final assignment
    (30_000, 32, 48, 2048, 40, 0.0, np.float32, "bf16x3"),  # BF16 final assignment
])
def test_minibatch_matches_oracle(n, d, k, batch, max_iter, tol, dtype, precision):
    X = blobs(n, d, k, seed=n + k + batch, dtype=dtype)
    X64 = X.astype(np.float64)
    init = O.kmeanspp(X64, k, 5)
    seed = 1234
    c_o, l_o, i_o, it_o = O.minibatch_run(X64, init, batch, max_iter, tol, seed)
    f = fit_mb(X, init, batch, max_iter, tol, seed, precision)
    assert f.n_iter() == it_o
    assert (bits(f.cluster_centers()) == bits(c_o)).all()
    assert (f.labels() == l_o).all()
    assert bits(np.array([f.inertia()]))[0] == bits(np.array([i_o]))[0]


def test_minibatch_n_init_keeps_strictly_smaller_inertia():
    X = blobs(8_000, 6, 10, seed=21, spread=2.5)
    inits = np.stack([O.kmeanspp(X, 10, s) for s in (1, 2, 3)])
    base = 77
    runs = [O.minibatch_run(X, inits[r], 512, 30, 0.0, base + r * 1000003) for r in range(3)]
    best = 0
    for r in range(1, 3):
        if runs[r][2] < runs[best][2]:
            best = r
    f = fit_mb(X, inits, 512, 30, 0.0, base, n_init=3)
    c_o, l_o, i_o, it_o = runs[best]
    assert (bits(f.cluster_centers()) == bits(c_o)).all()
    assert (f.labels() == l_o).all() and f.n_iter() == it_o and f.inertia() == i_o


def test_minibatch_without_init_is_deterministic_and_consistent():
    X = blobs(10_000, 5, 12, seed=4)
    m = F.MiniBatchKMeans(12).with_random_state(9).with_max_iter(30)
    a, b = m.fit(X), m.fit(X)
    assert (bits(a.cluster_centers()) == bits(b.cluster_centers())).all()
    assert (a.labels() == b.labels()).all() and a.inertia() == b.inertia()
    lab, inertia = O.assign(X, a.cluster_centers())
    assert (a.labels() == lab).all()
    assert abs(a.inertia() - inertia) <= 1e-12 * inertia
    assert (a.predict(X) == lab).all()


def test_minibatch_fit_golden_vectors():
    # whole fits without given centres (k-means++ and batches from each run's
    # generator) against the pure-Python restatement of mini_batch_kmeans.rs
    from tests._golden import arr, golden
    for case in golden()["minibatch_fit"]:
        X = arr(case["X"])
        m = (F.MiniBatchKMeans(case["k"]).with_batch_size(case["batch"])
             .with_max_iter(case["max_iter"]).with_tol(case["tol"]).with_n_init(case["n_init"])
             .with_random_state(case["seed"]))
        f = m.fit(X)
        assert f.n_iter() == case["n_iter"]
        assert list(f.labels()) == case["labels"]
        assert (bits(f.cluster_centers()) == bits(arr(case["centers"]))).all()
        assert f.inertia() == float.fromhex(case["inertia"])
