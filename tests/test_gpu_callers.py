"""Callers of the squared-Euclidean path on the GPU: pairwise_distances_argmin
(_min) (ferrolearn-metrics/src/pairwise.rs:564-629) and SpectralClustering's
K-Means step (ferrolearn-cluster/src/spectral.rs:283-291), against the
oracle bit for bit."""
import numpy as np
import pytest
import torch

import ferrolearn_amd as F
from oracle import oracle as O
from tests._golden import bits

pytestmark = pytest.mark.gpu


def blobs(n, d, k, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, size=(k, d))
    lab = rng.integers(0, k, size=n)
    return c[lab] + rng.normal(scale=spread, size=(n, d))


# ---------------------------------------------------------------- pairwise --

@pytest.mark.parametrize("shape", [(3000, 700, 37), (257, 31, 5), (1, 1, 1), (600, 65, 70)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pairwise_argmin_matches_oracle(shape, dtype):
    n, m, d = shape
    rng = np.random.default_rng(n + m + d)
    X = blobs(n, d, 12, 1).astype(dtype)
    Y = (blobs(m, d, 12, 2) + rng.normal(scale=0.1, size=(m, d))).astype(dtype)
    idx_o, min_o = O.pairwise_argmin_min(X, Y)
    idx, mins = F.pairwise_distances_argmin_min(X, Y)
    assert (idx == idx_o).all() and (bits(mins) == bits(min_o)).all()
    assert (F.pairwise_distances_argmin(X, Y) == idx_o).all()
    di, dm = F.pairwise_distances_argmin_min(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda())
    assert (di.cpu().numpy() == idx_o).all() and (bits(dm.cpu().numpy()) == bits(min_o)).all()


def test_pairwise_argmin_ties_cancellation_nan():
    rng = np.random.default_rng(4)
    # integer data with duplicated Y rows: equal distances, lowest index
    X = rng.integers(-4, 5, size=(2000, 6)).astype(np.float64)
    Y = rng.integers(-4, 5, size=(80, 6)).astype(np.float64)
    Y[50:60] = Y[3:13]
    for a, b in ((X, Y),
                 # far from the origin, where the expansion cancels
                 (1e8 + rng.normal(size=(500, 3)), 1e8 + rng.normal(size=(70, 3)))):
        idx_o, min_o = O.pairwise_argmin_min(a, b)
        idx, mins = F.pairwise_distances_argmin_min(a, b)
        assert (idx == idx_o).all() and (bits(mins) == bits(min_o)).all()
    Xn = rng.normal(size=(40, 4))
    Xn[3, 1] = np.nan
    Yn = rng.normal(size=(9, 4))
    Yn[5, 0] = np.nan
    idx_o, min_o = O.pairwise_argmin_min(Xn, Yn)
    idx, mins = F.pairwise_distances_argmin_min(Xn, Yn)
    assert (idx == idx_o).all() and (bits(mins) == bits(min_o)).all()


@pytest.mark.parametrize("shape", [(700, 300, 37), (257, 31, 5), (1, 1, 1), (600, 65, 70)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_euclidean_distances_matches_oracle(shape, dtype):
    n, m, d = shape
    rng = np.random.default_rng(n * 3 + m + d)
    X = blobs(n, d, 12, 5).astype(dtype)
    Y = (blobs(m, d, 12, 6) + rng.normal(scale=0.1, size=(m, d))).astype(dtype)
    want = O.euclidean_distances(X, Y)
    got = F.euclidean_distances(X, Y)
    assert got.shape == (n, m) and (bits(got) == bits(want)).all()
    gd = F.euclidean_distances(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda())
    assert (bits(gd.cpu().numpy()) == bits(want)).all()


def test_euclidean_distances_cancellation_nan_doc_example():
    d = F.euclidean_distances(np.array([[0.0, 0.0], [1.0, 0.0]]),
                              np.array([[0.0, 0.0], [0.0, 1.0]]))   # pairwise.rs:146-154
    assert d[0, 0] == 0.0 and d[0, 1] == 1.0 and d[1, 0] == 1.0 and d[1, 1] == np.sqrt(2.0)
    rng = np.random.default_rng(9)
    a = 1e8 + rng.normal(size=(300, 3))
    b = 1e8 + rng.normal(size=(70, 3))
    b[:5] = a[:5]
    a[7, 1] = np.nan
    assert (bits(F.euclidean_distances(a, b)) == bits(O.euclidean_distances(a, b))).all()


def test_pairwise_argmin_mixed_dtypes_widen():
    X = blobs(300, 8, 5, 3).astype(np.float32)
    Y = blobs(40, 8, 5, 4)
    idx_o, min_o = O.pairwise_argmin_min(X.astype(np.float64), Y)
    idx, mins = F.pairwise_distances_argmin_min(X, Y)
    assert (idx == idx_o).all() and (bits(mins) == bits(min_o)).all()


# ---------------------------------------------------------------- spectral --

def two_blobs():
    """spectral.rs:320-329."""
    return np.array([[0.0, 0.0], [0.2, 0.1], [-0.1, 0.2], [0.1, -0.1], [0.0, 0.1],
                     [10.0, 10.0], [10.2, 10.1], [9.9, 10.2], [10.1, 9.9], [10.0, 10.1]])


def test_spectral_reference_cases():
    x = two_blobs()
    lab = F.SpectralClustering(2).with_gamma(0.1).with_random_state(42).fit(x).labels()
    assert len(lab) == 10                                   # test_two_blobs_two_clusters
    assert len(set(lab[:5])) == 1 and len(set(lab[5:])) == 1 and lab[0] != lab[5]
    lab = F.SpectralClustering(2).with_random_state(1).fit(x).labels()
    assert ((lab >= 0) & (lab < 2)).all()                   # test_labels_in_valid_range
    assert (F.SpectralClustering(1).with_random_state(0).fit(x).labels() == 0).all()
    x3 = np.array([[0.0, 0.0], [5.0, 5.0], [10.0, 10.0]])  # test_n_clusters_equals_n_samples
    assert len(F.SpectralClustering(3).with_random_state(0).fit(x3).labels()) == 3
    x32 = np.array([[0.0, 0.0], [0.1, 0.1], [-0.1, 0.1], [10.0, 10.0], [10.1, 10.1],
                    [9.9, 10.1]], np.float32)               # test_f32_support
    assert len(F.SpectralClustering(2).with_gamma(0.1).with_random_state(42).fit(x32)
               .labels()) == 6
    m = F.SpectralClustering(2).with_gamma(0.1).with_random_state(7)   # reproducibility
    assert (m.fit(x).labels() == m.fit(x).labels()).all()
    assert (m.fit_predict(x) == m.fit(x).labels()).all()


@pytest.mark.parametrize("seed,n_init", [(None, 3), (11, 4)])
def test_spectral_kmeans_step_matches_oracle(seed, n_init):
    # step 5 given the embedding: n_init k-means++ restarts from the
    # reference's StdRng (seeds base + r), the strictly smallest inertia, then
    # predict on the embedding
    x = blobs(400, 5, 4, 8, spread=1.5)
    sc = F.SpectralClustering(4).with_gamma(0.05).with_n_init(n_init)
    if seed is not None:
        sc = sc.with_random_state(seed)
    e = sc.embed(x)
    base = 0 if seed is None else seed
    c_o, _, _, _ = O.kmeans_fit_stdrng(e, 4, 300, 1e-4, n_init, base)
    want = O.assign(e, c_o)[0]
    assert (sc.fit(x).labels() == want).all()
