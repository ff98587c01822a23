"""BisectingKMeans on the GPU (ferrolearn-cluster/src/bisecting_kmeans.rs):
labels, centres and inertia bit for bit against the golden runs of the
pure-Python restatement and the oracle, with both strategies."""
import numpy as np
import pytest

import ferrolearn_amd as F
from oracle import oracle as O
from tests._golden import arr, bits, golden

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STRAT = {"size": F.BisectingStrategy.LargestCluster, "sse": F.BisectingStrategy.LargestSSE}


def fit(X, n_clusters, max_iter, n_init, seed, strategy):
    m = (F.BisectingKMeans(n_clusters).with_max_iter(max_iter).with_n_init(n_init)
         .with_random_state(seed).with_bisecting_strategy(STRAT[strategy]))
    return m.fit(X)


def test_golden_runs_bit_exact():
    for run in golden()["bisecting"]:
        X = arr(run["X"])
        f = fit(X, run["n_clusters"], run["max_iter"], run["n_init"], int(run["seed"]),
                run["strategy"])
        assert list(f.labels()) == run["labels"]
        assert (bits(f.cluster_centers()) == bits(arr(run["centers"]))).all()
        assert f.inertia().hex() == run["inertia"]
        g = fit(torch.from_numpy(X).cuda(), run["n_clusters"], run["max_iter"], run["n_init"],
                int(run["seed"]), run["strategy"])
        assert list(g.labels().cpu().numpy()) == run["labels"]
        assert (bits(g.cluster_centers().cpu().numpy()) == bits(arr(run["centers"]))).all()
        assert g.inertia() == f.inertia()


def blobs(n, d, k, seed, spread, dtype):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-10, 10, size=(k, d))
    lab = rng.integers(0, k, size=n)
    return (cen[lab] + rng.normal(scale=spread, size=(n, d))).astype(dtype)


@pytest.mark.parametrize("n,d,k,n_init,strategy,dtype", [
    (20_000, 16, 8, 3, "size", np.float64),
    (15_000, 8, 6, 2, "sse", np.float32),
    (3_000, 64, 10, 10, "size", np.float32),
    (5_000, 3, 12, 1, "sse", np.float64),
])
def test_matches_oracle(n, d, k, n_init, strategy, dtype):
    X = blobs(n, d, k + 2, seed=n + d + k, spread=1.5, dtype=dtype)
    C_o, L_o, i_o = O.bisecting_kmeans(X.astype(np.float64), k, 300, n_init, 11, strategy)
    f = fit(X, k, 300, n_init, 11, strategy)
    assert f.n_clusters() == C_o.shape[0]
    assert (f.labels() == L_o).all()
    assert (bits(f.cluster_centers()) == bits(C_o)).all()
    assert f.inertia() == i_o
    # Predict / Transform (bisecting_kmeans.rs:531-605)
    assert (f.predict(X) == O.assign(X.astype(np.float64), C_o)[0]).all()
    assert (bits(f.transform(X[:500])) == bits(O.transform(X[:500].astype(np.float64), C_o))).all()


def test_reference_unit_tests():
    # bisecting_kmeans.rs:612-700
    x = np.array([[0.0, 0.0], [0.1, 0.1], [-0.1, 0.1], [10.0, 10.0], [10.1, 10.1], [9.9, 10.1],
                  [0.0, 10.0], [0.1, 10.1], [-0.1, 9.9]])
    f = F.BisectingKMeans(3).with_random_state(42).fit(x)
    L = f.labels()
    assert L[0] == L[1] == L[2] and L[3] == L[4] == L[5] and L[6] == L[7] == L[8]
    assert len({L[0], L[3], L[6]}) == 3
    assert (f.predict(x) == L).all()
    assert (F.BisectingKMeans(3).with_random_state(42).fit_predict(x) == L).all()
    one = F.BisectingKMeans(1).with_random_state(42).fit(
        np.array([[1.0, 1.0], [2.0, 2.0], [3.0, 3.0], [4.0, 4.0]]))
    assert (one.labels() == 0).all() and one.n_clusters() == 1 and one.inertia() == 10.0
    with pytest.raises(F.ShapeMismatch, match="fitted BisectingKMeans model"):
        f.predict(np.ones((2, 3)))


def test_identical_samples_stop_with_an_error():
    # the reference keeps drawing seeds for a cluster it cannot split
    with pytest.raises(F.NumericalInstability, match="left one side empty"):
        F.BisectingKMeans(2).with_n_init(1).fit(np.ones((10, 2)))
