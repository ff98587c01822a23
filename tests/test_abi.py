"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and validates arguments on the host in the reference's
order (kmeans.rs:421-449, knn.rs:277-301) before touching the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import ferrolearn_amd as F
from ferrolearn_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "ferrolearn_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_version_and_device_count_never_fail():
    assert L.lib().fl_version() >= 100
    assert L.lib().fl_device_count() >= 0


def test_kmeans_validation_order():
    x = np.ones((3, 2))
    with pytest.raises(F.InvalidParameter, match="n_clusters"):
        F.KMeans(0).fit(x)
    with pytest.raises(F.InsufficientSamples, match="actual 0"):
        F.KMeans(2).fit(np.zeros((0, 2)))
    with pytest.raises(F.InsufficientSamples, match="required 4, actual 3"):
        F.KMeans(4).fit(x)
    with pytest.raises(F.InvalidParameter, match="n_init"):
        F.KMeans(2).with_n_init(0).fit(x)
    # k == 0 is reported before the empty input (kmeans.rs:424 before :431)
    with pytest.raises(F.InvalidParameter):
        F.KMeans(0).fit(np.zeros((0, 2)))


def test_minibatch_validation_order():
    # mini_batch_kmeans.rs:495-526: n_clusters, batch_size, n_init, then the
    # sample count -- each check fires before the later ones.
    x = np.ones((3, 2))
    empty = np.zeros((0, 2))
    with pytest.raises(F.InvalidParameter, match="n_clusters"):
        F.MiniBatchKMeans(0).with_batch_size(0).with_n_init(0).fit(empty)
    with pytest.raises(F.InvalidParameter, match="batch_size"):
        F.MiniBatchKMeans(2).with_batch_size(0).with_n_init(0).fit(empty)
    with pytest.raises(F.InvalidParameter, match="n_init"):
        F.MiniBatchKMeans(2).with_n_init(0).fit(empty)
    with pytest.raises(F.InsufficientSamples, match="actual 0"):
        F.MiniBatchKMeans(2).fit(empty)
    with pytest.raises(F.InsufficientSamples, match="required 4, actual 3"):
        F.MiniBatchKMeans(4).fit(x)


def test_minibatch_defaults():
    m = F.MiniBatchKMeans(5)
    assert (m.batch_size, m.max_iter, m.tol, m.n_init, m.random_state) == (1024, 100, 0.0, 3, None)


def test_kmeans_predict_shape_mismatch():
    fitted = F.FittedKMeans(np.zeros((2, 2)), np.zeros(4, np.int64), 0.0, 1, {})
    with pytest.raises(F.ShapeMismatch, match="expected \\[2\\], actual \\[3\\]"):
        fitted.predict(np.ones((2, 3)))
    with pytest.raises(F.ShapeMismatch):
        fitted.transform(np.ones((2, 3)))


def test_knn_validation_order():
    x = np.ones((3, 2))
    with pytest.raises(F.ShapeMismatch):
        F.KNeighborsClassifier().fit(x, [0, 1])
    with pytest.raises(F.InvalidParameter, match="n_neighbors"):
        F.KNeighborsClassifier().with_n_neighbors(0).fit(x, [0, 1, 2])
    with pytest.raises(F.InsufficientSamples, match="required 5, actual 3"):
        F.KNeighborsClassifier().fit(x, [0, 1, 2])
    # y length is checked before n_neighbors (knn.rs:280 before :288)
    with pytest.raises(F.ShapeMismatch):
        F.KNeighborsRegressor().with_n_neighbors(0).fit(x, [0.0])
    with pytest.raises(F.InsufficientSamples):
        F.KNeighborsRegressor().fit(x, [0.0, 1.0, 2.0])


def test_knn_large_k_is_only_bounded_by_n():
    # knn.rs:288-301: any 1 <= n_neighbors <= n is accepted; k > n is the
    # reference's InsufficientSamples, raised before any device work.
    x = np.ones((100, 2))
    with pytest.raises(F.InsufficientSamples, match="required 101, actual 100"):
        F.KNeighborsClassifier().with_n_neighbors(101).fit(x, np.zeros(100))


def test_gemm_shape_mismatch():
    a = np.ones((2, 3))
    b = np.ones((4, 2))
    st = L.lib().fl_backend_gemm_f64(a.ctypes.data_as(C.c_void_p), 2, 3,
                                     b.ctypes.data_as(C.c_void_p), 4, 2,
                                     np.zeros(4).ctypes.data_as(C.c_void_p))
    assert st == L.FL_SHAPE_MISMATCH
    # the reference's fields: expected = A's shape, actual = B's shape
    assert "expected [2, 3], actual [4, 2]" in L.last_error()
    assert "gemm: A is 2x3 but B is 4x2 (inner dimensions 3 != 4)" in L.last_error()


def test_no_device_gives_device_error_not_a_crash():
    if L.lib().fl_device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(F.DeviceError):
        F.KMeans(2).with_n_init(1).fit(np.ones((4, 2)))


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "ferrolearn_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")) or f == "Makefile":
                txt = open(os.path.join(dirpath, f), errors="ignore").read()
                assert not re.search(r"(from|import)\s+oracle|liboracle|oracle/|ref_cpu", txt), f


STRUCTS = {
    "fl_report": L.fl_report,
    "fl_kmeans_params": L.fl_kmeans_params,
    "fl_minibatch_params": L.fl_minibatch_params,
    "fl_router_params": L.fl_router_params,
    "fl_router_train_params": L.fl_router_train_params,
    "fl_router_report": L.fl_router_report,
    "fl_kmeans_problem": L.fl_kmeans_problem,
    "fl_bisecting_params": L.fl_bisecting_params,
}


def test_ctypes_structs_match_the_header(tmp_path):
    # sizeof and offsetof of every field, as a C compiler lays them out
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "ferrolearn_hip.h"',
             "int main(void) {"]
    for name, cls in STRUCTS.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src),
                           "-o", str(exe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for name, cls in STRUCTS.items():
        assert int(got[name]) == C.sizeof(cls), name
        for field, _ in cls._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(cls, field).offset, (name, field)


def test_router_defaults_and_null_arguments():
    p = L.fl_router_params()
    L.lib().fl_router_params_default(C.byref(p), 4096)
    assert (p.n_cells, p.cell_iter, p.hidden, p.seed) == (4096, 10, 512, 0)
    t = L.fl_router_train_params()
    L.lib().fl_router_train_params_default(C.byref(t))
    assert (t.n_neighbors, t.epochs, t.batch_size, t.target) == (10, 5, 4096, L.FL_TARGET_CERTIFY)
    h = C.c_void_p()
    rep = L.fl_router_report()
    assert L.lib().fl_router_build(None, C.byref(p), C.byref(h), C.byref(rep)) == \
        L.FL_INVALID_PARAMETER
    assert "knn" in L.last_error()
    assert L.lib().fl_router_kneighbors_device(None, None, 0, 0, 4, 0, 2, 0, None, None,
                                               None) == L.FL_INVALID_PARAMETER
    assert L.lib().fl_router_n_params(None) == 0
    L.lib().fl_router_free(None)
    loss = C.c_double()
    assert L.lib().fl_router_mlp_gradients(4, 64, 8, None, None, 1, None, 1, 0.0,
                                           C.byref(loss), None) == L.FL_INVALID_PARAMETER


def test_nearest_neighbors_fit_validation():
    # nearest_neighbors.rs:203-222: n_neighbors == 0, then n < n_neighbors
    x = np.ones((2, 2))
    with pytest.raises(F.InvalidParameter, match="n_neighbors"):
        F.NearestNeighbors().with_n_neighbors(0).fit(x)
    with pytest.raises(F.InsufficientSamples,
                       match="NearestNeighbors requires at least n_neighbors=5 samples"):
        F.NearestNeighbors().with_n_neighbors(5).fit(x)


def test_pairwise_argmin_validation_order():
    # pairwise.rs:77-117: feature dimensions first, then empty inputs
    with pytest.raises(F.ShapeMismatch, match="feature dimensions must match"):
        F.pairwise_distances_argmin(np.ones((0, 2)), np.ones((3, 3)))
    with pytest.raises(F.InsufficientSamples, match="at least one row"):
        F.pairwise_distances_argmin(np.ones((0, 2)), np.ones((3, 2)))
    with pytest.raises(F.InsufficientSamples, match="feature dimension must be at least 1"):
        F.pairwise_distances_argmin_min(np.ones((2, 0)), np.ones((3, 0)))
    with pytest.raises(F.InvalidParameter, match="metric"):
        F.pairwise_distances_argmin(np.ones((2, 2)), np.ones((3, 2)), F.Metric.Manhattan)


def test_euclidean_distances_validation_order():
    # pairwise.rs:163-164 (check_feature_dim, then check_non_empty)
    with pytest.raises(F.ShapeMismatch, match="feature dimensions must match"):
        F.euclidean_distances(np.ones((0, 2)), np.ones((3, 3)))
    with pytest.raises(F.InsufficientSamples, match="at least one row"):
        F.euclidean_distances(np.ones((2, 2)), np.ones((0, 2)))
    with pytest.raises(F.InsufficientSamples, match="feature dimension must be at least 1"):
        F.euclidean_distances(np.ones((2, 0)), np.ones((3, 0)))


def test_bisecting_defaults_and_validation_order():
    # bisecting_kmeans.rs:93-102 defaults; :326-366 checks in order
    p = L.fl_bisecting_params()
    L.lib().fl_bisecting_params_default(C.byref(p), 5)
    assert (p.n_clusters, p.max_iter, p.n_init, p.has_random_state, p.strategy) == \
        (5, 300, 10, 0, 0)
    x = np.ones((3, 2))
    with pytest.raises(F.InvalidParameter, match="n_clusters"):
        F.BisectingKMeans(0).with_n_init(0).fit(x)
    with pytest.raises(F.InvalidParameter, match="n_init"):
        F.BisectingKMeans(4).with_n_init(0).fit(x)
    with pytest.raises(F.InsufficientSamples,
                       match="required 4, actual 3: BisectingKMeans requires at least"):
        F.BisectingKMeans(4).fit(x)
    with pytest.raises(F.InsufficientSamples, match="required 2, actual 0"):
        F.BisectingKMeans(2).fit(np.ones((0, 2)))


def test_spectral_validation():
    # spectral.rs:240-258
    x = np.ones((10, 2))
    with pytest.raises(F.InvalidParameter, match="n_clusters"):
        F.SpectralClustering(0).fit(x)
    for g in (0.0, -1.0):
        with pytest.raises(F.InvalidParameter, match="gamma"):
            F.SpectralClustering(2).with_gamma(g).fit(x)
    with pytest.raises(F.InsufficientSamples):
        F.SpectralClustering(2).fit(np.zeros((0, 2)))
    with pytest.raises(F.InsufficientSamples):
        F.SpectralClustering(3).fit(np.zeros((1, 2)))


def test_rs_surface_errors_and_defaults():
    # ferrolearn-python: constructor defaults (clusterers.rs:25-26,
    # classifiers.rs:306-307, extras.rs:700-701), RuntimeError before fit,
    # ValueError with FerroError's Display text (error.rs:34-76)
    from ferrolearn_amd import _ferrolearn_rs as R
    km = R._RsKMeans()
    assert (km.n_clusters, km.max_iter, km.tol, km.n_init, km.random_state) == \
        (8, 300, 1e-4, 10, None)
    assert R._RsKNeighborsClassifier().n_neighbors == 5
    assert R._RsKNeighborsRegressor().n_neighbors == 5
    mb = R._RsMiniBatchKMeans()
    assert (mb.n_clusters, mb.max_iter, mb.random_state) == (8, 100, None)
    for obj, call in ((km, lambda: km.predict(np.ones((2, 2)))),
                      (km, lambda: km.labels_),
                      (mb, lambda: mb.labels_)):
        with pytest.raises(RuntimeError, match="not fitted"):
            call()
    with pytest.raises(ValueError, match=r"^Invalid parameter `n_clusters`: must be at least 1$"):
        R._RsKMeans(n_clusters=0).fit(np.ones((5, 2)))
    with pytest.raises(ValueError,
                       match=r"^Insufficient samples: need at least 5, got 3 \(KNeighborsClassifier"):
        R._RsKNeighborsClassifier().fit(np.ones((3, 2)), np.zeros(3))
    with pytest.raises(ValueError, match=r"^Shape mismatch in y length must match"):
        R._RsKNeighborsRegressor().fit(np.ones((6, 2)), np.zeros(4))
