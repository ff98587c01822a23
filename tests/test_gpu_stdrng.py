"""KMeans / MiniBatchKMeans drawing from the reference's StdRng stream
(with_rng("stdrng")): k-means++ with the reference's plain left folds
(kmeans.rs:169-253) and whole seeded fits, against the oracle and the
pure-Python golden vectors, bit for bit.  The stream is ChaCha12 keyed as
rand 0.9.4's StdRng::seed_from_u64 (its ChaCha block is pinned to RFC 8439's
vectors in test_oracle.py; the seeding and range sampling are restated from
the crates' published algorithms and pinned against the golden restatement)."""
import numpy as np
import pytest

import ferrolearn_amd as F
from oracle import oracle as O
from tests._golden import arr, bits, golden

pytestmark = pytest.mark.gpu

G = golden()


def blobs(n, d, k, seed, spread=1.0, lo=-10.0, hi=10.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, size=(k, d))
    lab = rng.integers(0, k, size=n)
    return c[lab] + rng.normal(scale=spread, size=(n, d))


def test_kmeanspp_stdrng_golden_vectors():
    # max_iter = 0 returns the k-means++ centres of the (single) run
    for case in G["kmeanspp_stdrng"]:
        X = arr(case["X"])
        f = (F.KMeans(case["k"]).with_n_init(1).with_max_iter(0).with_rng("stdrng")
             .with_random_state(int(case["seed"])).fit(X))
        assert (bits(f.cluster_centers()) == bits(X[case["picks"]])).all(), (X.shape, case["k"])


def test_kmeans_fit_stdrng_golden_vectors():
    for case in G["kmeans_fit_stdrng"]:
        X = arr(case["X"])
        f = (F.KMeans(case["k"]).with_n_init(case["n_init"]).with_max_iter(case["max_iter"])
             .with_tol(case["tol"]).with_rng("stdrng").with_random_state(case["seed"]).fit(X))
        assert f.n_iter() == case["n_iter"] and list(f.labels()) == case["labels"]
        assert (bits(f.cluster_centers()) == bits(arr(case["centers"]))).all()
        # the device sums the inertia in 1024-row blocks (within 1e-12 of the
        # reference's row-order fold)
        want = float.fromhex(case["inertia"])
        assert abs(f.inertia() - want) <= 1e-12 * want


def test_config1_kmeans_stdrng_matches_oracle():
    # KMeans Lloyd in f64 at the reference's CPU-runnable case: 100,000 x 32
    # blobs (16 centres in [-10, 10]^32, std 1), k = 16, k-means++ seeded,
    # n_init 1, max_iter 100, tol 1e-4
    X = blobs(100_000, 32, 16, 0)
    for seed in (0, 42):
        f = (F.KMeans(16).with_n_init(1).with_max_iter(100).with_tol(1e-4).with_rng("stdrng")
             .with_random_state(seed).fit(X))
        c_o, l_o, i_o, it_o = O.kmeans_fit_stdrng(X, 16, 100, 1e-4, 1, seed)
        assert f.n_iter() == it_o and (f.labels() == l_o).all(), seed
        assert (bits(f.cluster_centers()) == bits(c_o)).all()
        assert abs(f.inertia() - i_o) <= 1e-12 * i_o


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_kmeans_stdrng_n_init_and_filters(precision):
    X = blobs(20_000, 12, 40, 3, spread=1.5)
    f = (F.KMeans(40).with_n_init(3).with_max_iter(30).with_precision(precision)
         .with_rng("stdrng").with_random_state(9).fit(X))
    c_o, l_o, i_o, it_o = O.kmeans_fit_stdrng(X, 40, 30, 1e-4, 3, 9)
    assert f.n_iter() == it_o and (f.labels() == l_o).all()
    assert (bits(f.cluster_centers()) == bits(c_o)).all()


def test_minibatch_stdrng_matches_oracle_and_golden():
    for case in G["minibatch_fit_stdrng"]:
        X = arr(case["X"])
        f = (F.MiniBatchKMeans(case["k"]).with_batch_size(case["batch"])
             .with_max_iter(case["max_iter"]).with_tol(case["tol"]).with_n_init(case["n_init"])
             .with_rng("stdrng").with_random_state(case["seed"]).fit(X))
        assert f.n_iter() == case["n_iter"] and list(f.labels()) == case["labels"]
        assert (bits(f.cluster_centers()) == bits(arr(case["centers"]))).all()
    X = blobs(30_000, 8, 10, 4)
    f = (F.MiniBatchKMeans(10).with_max_iter(40).with_rng("stdrng").with_random_state(3).fit(X))
    c_o, l_o, i_o, it_o = O.minibatch_fit_stdrng(X, 10, 1024, 40, 0.0, 3, 3)
    assert f.n_iter() == it_o and (f.labels() == l_o).all()
    assert (bits(f.cluster_centers()) == bits(c_o)).all()
