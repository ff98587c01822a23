"""Cell router parity on the GPU.

Every routed answer must equal the brute force's (distances and indices bit
for bit, (distance, index) order), whatever the route, n_probe or training
state; the network's gradients must match the numpy restatement of mlp.rs
(oracle.mlp_loss_grad with the same bfloat16 roundings); and a trained
network must certify more queries than the untrained one."""
import numpy as np
import pytest

import ferrolearn_amd as F
from ferrolearn_amd.router import mlp_gradients
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def mixture(N, d, comps, seed, spread=10.0, dtype=np.float32):
    """N points of one fixed Gaussian mixture (means and spreads from seed
    0); `seed` only draws the points, so database and queries share it."""
    g = np.random.default_rng(0)
    means = g.uniform(-spread, spread, size=(comps, d))
    sd = g.uniform(0.5, 1.5, size=comps)
    rng = np.random.default_rng(seed)
    c = rng.integers(0, comps, size=N)
    X = means[c] + rng.normal(size=(N, d)) * sd[c, None]
    return X.astype(dtype), c


@pytest.fixture(scope="module")
def setup():
    X, comp = mixture(20000, 24, 40, seed=1)
    # exact duplicates of some rows: distance ties resolved by index
    X = np.concatenate([X, X[:500]], axis=0)
    y = np.arange(X.shape[0]) % 7
    fitted = (F.KNeighborsClassifier().with_n_neighbors(10).with_precision("bf16x3")
              .fit(X, y))
    router = F.CellRouter(fitted, n_cells=40, cell_iter=8, hidden=64, seed=1)
    Qt, _ = mixture(4000, 24, 40, seed=2)
    untrained = router.params()
    rep = router.train(Qt, n_neighbors=10, epochs=10, batch_size=512, learning_rate=3e-3, seed=3)
    Q, _ = mixture(2000, 24, 40, seed=4)
    return X, fitted, router, untrained, rep, Q


def test_cells_cover_rows(setup):
    X, fitted, router, *_ = setup
    cen, rad, cnt, lab = router.cells()
    assert cnt.sum() == X.shape[0]
    assert (np.bincount(lab, minlength=router.n_cells) == cnt).all()
    # every row lies within its cell's radius (exact folds on the CPU)
    sample = np.arange(0, X.shape[0], 37)
    for i in sample:
        c = lab[i]
        assert O.euclid(X[i].astype(np.float64), cen[c]) <= rad[c]


def test_training_lowers_loss(setup):
    rep = setup[4]
    assert rep["loss_last"] < rep["loss_first"]


@pytest.mark.parametrize("route", ["network", "bounds", "centres"])
@pytest.mark.parametrize("n_probe", [1, 4])
def test_routed_equals_brute_force(setup, route, n_probe):
    X, fitted, router, _, _, Q = setup
    bd, bi = fitted.kneighbors(Q)
    rd, ri = router.kneighbors(Q, n_probe=n_probe, route=route)
    assert (ri == bi).all()
    assert (rd.view(np.int64) == bd.view(np.int64)).all()
    rep = router.report
    assert 0 <= rep["n_fallback"] <= Q.shape[0]
    assert rep["rows_scanned"] > 0


def test_routed_against_oracle(setup):
    X, fitted, router, _, _, Q = setup
    rd, ri = router.kneighbors(Q[:40], n_probe=4)
    od, oi = O.kneighbors(X, Q[:40], 10)
    assert (ri == oi).all()
    assert (rd.view(np.int64) == od.view(np.int64)).all()


@pytest.mark.parametrize("k", [1, 17, 32])
def test_routed_k_and_dtypes(setup, k):
    X, fitted, router, _, _, Q = setup
    Q64 = Q[:500].astype(np.float64)
    bd, bi = fitted.kneighbors(Q64, n_neighbors=k)
    rd, ri = router.kneighbors(Q64, n_neighbors=k, n_probe=3)
    assert (ri == bi).all()
    assert (rd.view(np.int64) == bd.view(np.int64)).all()


def test_all_cells_always_certify(setup):
    X, fitted, router, _, _, Q = setup
    rd, ri = router.kneighbors(Q[:300], n_probe=min(31, router.n_cells), route="bounds")
    # 31 of 40 cells scanned, smallest bounds first: every query certifies
    assert router.report["n_rescan"] == 0 and router.report["n_fallback"] == 0
    bd, bi = fitted.kneighbors(Q[:300])
    assert (ri == bi).all()
    assert (rd.view(np.int64) == bd.view(np.int64)).all()


def test_network_gain_over_untrained(setup):
    X, fitted, router, untrained, _, Q = setup
    # n_rescan: queries whose certificate failed after the n_probe cells the
    # route chose (they are searched again, then by brute force)
    router.kneighbors(Q, n_probe=2, route="network")
    trained_fb = router.report["n_rescan"]
    trained_rows = router.report["rows_scanned"]
    assert router.report["n_fallback"] <= trained_fb
    trained = router.params()
    router.set_params(untrained)
    try:
        router.kneighbors(Q, n_probe=2, route="network")
        untrained_fb = router.report["n_rescan"]
    finally:
        router.set_params(trained)
    router.kneighbors(Q, n_probe=2, route="bounds")
    bounds_fb = router.report["n_rescan"]
    print(f"fallback at n_probe=2: trained {trained_fb}, untrained {untrained_fb}, "
          f"bounds {bounds_fb} of {Q.shape[0]}; rows scanned {trained_rows}")
    assert trained_fb < untrained_fb
    # the certificate covers most queries once the right cells are scanned
    assert bounds_fb < 0.5 * Q.shape[0]
    # the routed search (with its rescans) scans a small part of the database
    assert trained_rows < 0.25 * X.shape[0] * Q.shape[0]


@pytest.mark.parametrize("alpha", [0.0, 1e-3])
def test_gradients_match_oracle(alpha):
    rng = np.random.default_rng(7)
    d, h, c, B, k = 20, 64, 45, 130, 5
    n = d * h + h + h * h + h + h * c + c
    p = rng.normal(scale=0.2, size=n).astype(np.float32)
    X = rng.normal(size=(B, d)).astype(np.float32)
    cells = rng.integers(0, c, size=(B, k)).astype(np.int32)
    cells[::3, 3:] = -1          # rows with fewer targets
    cells[7, :] = -1             # a row without targets
    loss, g = mlp_gradients(p, X, cells, h, c, alpha=alpha)
    ol, og = O.mlp_loss_grad(p, X, cells, d, h, c, alpha=alpha, emulate_bf16=True)
    # fp32 accumulation against float64, same bfloat16 roundings: loss to
    # 1e-5 relative, every parameter block to 1e-2 in norm
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    o = 0
    for s in [(d, h), (h,), (h, h), (h,), (h, c), (c,)]:
        m = int(np.prod(s))
        gg, oo = g[o:o + m].astype(np.float64), og[o:o + m]
        assert np.linalg.norm(gg - oo) <= 1e-2 * np.linalg.norm(oo) + 1e-12, s
        o += m


def test_neighbor_target_training(setup):
    X, fitted, router, untrained, _, Q = setup
    r2 = F.CellRouter(fitted, n_cells=40, cell_iter=8, hidden=64, seed=1)
    Qt, _ = mixture(4000, 24, 40, seed=2)
    rep = r2.train(Qt, n_neighbors=10, epochs=10, batch_size=512, learning_rate=3e-3, seed=3,
                   target="neighbors")
    assert rep["loss_last"] < rep["loss_first"]
    rd, ri = r2.kneighbors(Q, n_probe=2)
    bd, bi = fitted.kneighbors(Q)
    assert (ri == bi).all()
    print(f"neighbors-target rescans at n_probe=2: {r2.report['n_rescan']}")


def test_router_validation(setup):
    X, fitted, router, *_ = setup
    with pytest.raises(F.InvalidParameter):
        router.kneighbors(X[:3], n_probe=0)
    with pytest.raises(F.InvalidParameter):
        router.kneighbors(X[:3], n_probe=40)
    with pytest.raises(F.ShapeMismatch):
        router.kneighbors(X[:3, :5])
    with pytest.raises(F.InvalidParameter):
        F.CellRouter(fitted, n_cells=16, hidden=100)


def test_unstructured_data_full_scan():
    # Uniform data: the bounds rarely certify, so queries go through the
    # widening rescans and then the all-cells scan; every answer is exact
    # and none needs the brute force.
    rng = np.random.default_rng(11)
    X = rng.uniform(-1, 1, size=(20000, 16)).astype(np.float32)
    fitted = F.KNeighborsRegressor().with_n_neighbors(10).fit(X, np.zeros(len(X)))
    r = F.CellRouter(fitted, n_cells=64, cell_iter=5, hidden=64, seed=2)
    Q = rng.uniform(-1, 1, size=(700, 16)).astype(np.float32)
    rd, ri = r.kneighbors(Q, n_probe=2, route="bounds")
    bd, bi = fitted.kneighbors(Q)
    assert (ri == bi).all()
    assert (rd.view(np.int64) == bd.view(np.int64)).all()
    rep = r.report
    print(f"unstructured: rescans {rep['n_rescan']}, brute force {rep['n_fallback']}")
    assert rep["n_rescan"] > 0 and rep["n_fallback"] == 0


def test_non_finite_queries(setup):
    # A query with an infinite feature never certifies: after the all-cells
    # pass it goes to the handle's brute force, so its answer (every
    # distance inf, ties by index) is the brute force's.  A NaN feature gives
    # NaN distances, which the reference rejects (kdtree.rs:386); both
    # searches raise the same error.
    X, fitted, router, _, _, Q = setup
    Qi = Q[:64].copy()
    Qi[10, 0] = np.inf
    rd, ri = router.kneighbors(Qi, n_probe=2)
    bd, bi = fitted.kneighbors(Qi)
    assert (ri == bi).all()
    assert (rd.view(np.int64) == bd.view(np.int64)).all()
    assert router.report["n_fallback"] == 1
    Qn = Q[:64].copy()
    Qn[3, 5] = np.nan
    with pytest.raises(F.NumericalInstability):
        fitted.kneighbors(Qn)
    with pytest.raises(F.NumericalInstability):
        router.kneighbors(Qn, n_probe=2)


def test_large_cells():
    # Three cells of about 13,000 rows, one probe, k up to 32: the first
    # pass, the bound-skipping rescans and the all-cells pass on large
    # cells.  Answers stay the brute force's.
    rng = np.random.default_rng(5)
    X = rng.normal(size=(40000, 12)).astype(np.float32)
    fitted = F.KNeighborsRegressor().with_n_neighbors(10).fit(X, np.zeros(len(X)))
    r = F.CellRouter(fitted, n_cells=3, cell_iter=4, hidden=64, seed=1)
    assert r.cells()[2].max() > 8192
    Q = rng.normal(size=(500, 12)).astype(np.float32)
    for route in ("bounds", "network"):
        for k in (10, 32):
            rd, ri = r.kneighbors(Q, n_neighbors=k, n_probe=1, route=route)
            bd, bi = fitted.kneighbors(Q, n_neighbors=k)
            assert (ri == bi).all(), (route, k)
            assert (rd.view(np.int64) == bd.view(np.int64)).all(), (route, k)


def aniso_mixture(N, comps, seed, d=96):
    """The learned-router workload's data family at reduced size: a mixture of
    `comps` anisotropic Gaussians in d = 96 (random rotations, axis std
    log-uniform in [0.2, 3], means uniform in [-10, 10]); the mixture is fixed
    by seed 0 and `seed` draws the points."""
    g = np.random.default_rng(0)
    means = g.uniform(-10, 10, size=(comps, d))
    sd = np.exp(g.uniform(np.log(0.2), np.log(3.0), size=(comps, d)))
    rot = np.linalg.qr(g.standard_normal((comps, d, d)))[0]
    rng = np.random.default_rng(seed)
    c = rng.integers(0, comps, size=N)
    z = rng.standard_normal((N, d)) * sd[c]
    X = np.empty((N, d))
    for j in range(comps):
        m = c == j
        X[m] = z[m] @ rot[j].T + means[j]
    return X.astype(np.float32)


@pytest.fixture(scope="module")
def config5_setup():
    X = aniso_mixture(200_000, 256, seed=1)
    fitted = (F.KNeighborsClassifier().with_n_neighbors(10).with_precision("bf16x3")
              .fit(X, np.arange(X.shape[0]) % 10))
    router = F.CellRouter(fitted, n_cells=256, cell_iter=8, hidden=512, seed=0)
    Qt = aniso_mixture(40_000, 256, seed=2)
    rep = router.train(Qt, n_neighbors=10, epochs=2, batch_size=2048, learning_rate=1e-3, seed=0)
    Q = aniso_mixture(5000, 256, seed=3)
    return X, fitted, router, rep, Q


@pytest.mark.parametrize("route", ["network", "bounds", "centres"])
@pytest.mark.parametrize("n_probe", [2, 4])
def test_config5_shape_routed_equals_brute_force(config5_setup, route, n_probe):
    # d = 96, 256 cells over 200,000 rows, a 96-512-512-256 network: every
    # routed answer is the brute force's, bit for bit.
    X, fitted, router, rep, Q = config5_setup
    assert rep["loss_last"] < rep["loss_first"]
    bd, bi = fitted.kneighbors(Q)
    rd, ri = router.kneighbors(Q, n_probe=n_probe, route=route)
    assert (ri == bi).all()
    assert (rd.view(np.int64) == bd.view(np.int64)).all()
    print(f"d=96 {route} n_probe={n_probe}: rescans {router.report['n_rescan']}, "
          f"brute force {router.report['n_fallback']} of {Q.shape[0]}")


def test_config5_shape_against_oracle(config5_setup):
    X, fitted, router, rep, Q = config5_setup
    rd, ri = router.kneighbors(Q[:16], n_probe=2)
    od, oi = O.kneighbors(X, Q[:16], 10)
    assert (ri == oi).all()
    assert (rd.view(np.int64) == od.view(np.int64)).all()


def test_gradients_match_oracle_hidden512():
    # the router network's widths at the workload's hidden size
    rng = np.random.default_rng(8)
    d, h, c, B, k = 96, 512, 256, 64, 10
    n = d * h + h + h * h + h + h * c + c
    p = (rng.normal(size=n) * 0.05).astype(np.float32)
    X = aniso_mixture(B, 256, seed=9)
    cells = rng.integers(0, c, size=(B, k)).astype(np.int32)
    cells[::5, 6:] = -1
    loss, g = mlp_gradients(p, X, cells, h, c)
    ol, og = O.mlp_loss_grad(p, X, cells, d, h, c, emulate_bf16=True)
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    o = 0
    for s in [(d, h), (h,), (h, h), (h,), (h, c), (c,)]:
        m = int(np.prod(s))
        gg, oo = g[o:o + m].astype(np.float64), og[o:o + m]
        assert np.linalg.norm(gg - oo) <= 1e-2 * np.linalg.norm(oo) + 1e-12, s
        o += m


def test_config5_shape_routes_compared(config5_setup):
    # The same search with three probe orders at equal n_probe: the trained
    # network, the smallest lower bounds, and the nearest centres (the
    # classic inverted-file order, no network).  Every answer is the brute
    # force's; the counts of queries whose first certificate failed are
    # printed side by side.
    X, fitted, router, rep, Q = config5_setup
    bd, bi = fitted.kneighbors(Q)
    rates = {}
    for n_probe in (1, 2, 4):
        for route in ("network", "bounds", "centres"):
            rd, ri = router.kneighbors(Q, n_probe=n_probe, route=route)
            assert (ri == bi).all() and (rd.view(np.int64) == bd.view(np.int64)).all()
            rates[(route, n_probe)] = router.report["n_rescan"]
    print("first-certificate failures of", Q.shape[0], "queries:",
          {f"{r}@{p}": v for (r, p), v in rates.items()})
    for n_probe in (1, 2, 4):
        # On these mixtures the cells a certified search needs are, for
        # nearly every query, the cells with the nearest centres: the trained
        # network matches the nearest-centre order to within 1 % of the
        # queries and does not beat it.
        assert rates[("network", n_probe)] <= rates[("centres", n_probe)] + Q.shape[0] // 100


def _per_cell(keys, ids):
    return [dict(zip(i.tolist(), k.tolist())) for k, i in zip(keys, ids)]


def test_mfma_bounds_are_below_exact_bounds(setup):
    # Every cell's bound from the BF16x3 MFMA lower bound of the fold is at
    # most the bound from the exact fold, and close to it; with kb = n_cells
    # every cell is compared.
    X, fitted, _, _, _, Q = setup
    r = F.CellRouter(fitted, n_cells=24, cell_iter=5, hidden=64, seed=3)
    for q in (Q[:500], Q[:500].astype(np.float64)):
        km, im = r.cell_bounds(q, 24)
        ke, ie = r.cell_bounds(q, 24, exact=True)
        gaps = []
        for dm, de in zip(_per_cell(km, im), _per_cell(ke, ie)):
            assert set(dm) == set(de) == set(range(24))
            for c, v in de.items():
                assert dm[c] <= v
                gaps.append((v - dm[c]) / (1.0 + abs(v)))
        # the BF16x3 error term is about 2e-4 |q||mu| in fold units
        assert np.median(gaps) < 1e-3 and max(gaps) < 0.1


def test_mfma_bounds_config5_shape(config5_setup):
    X, fitted, router, rep, Q = config5_setup
    km, im = router.cell_bounds(Q[:2000], 32)
    ke, ie = router.cell_bounds(Q[:2000], 32, exact=True)
    assert (km[:, -1] <= ke[:, -1]).all()
    same = 0
    for dm, de in zip(_per_cell(km, im), _per_cell(ke, ie)):
        for c in set(dm) & set(de):
            assert dm[c] <= de[c]
            same += 1
    assert same > 0.9 * 32 * 2000


def test_mfma_bounds_fall_back_outside_bf16_range(setup):
    # a query beyond 2^56 (or non-finite) sends the whole call to the exact
    # bounds, which then equal the exact path's keys
    X, fitted, router, *_ , Q = setup
    q = Q[:64].astype(np.float64).copy()
    q[3, 0] = 2.0 ** 60
    km, im = router.cell_bounds(q, 8)
    ke, ie = router.cell_bounds(q, 8, exact=True)
    assert (km.view(np.int64) == ke.view(np.int64)).all() and (im == ie).all()
