"""The allocation-free KMeans predict (fl_kmeans_predict_device_noalloc)
against fl_kmeans_predict, through the C ABI: labels bit for bit, on calls in
a row (the second and later with smaller inputs, so the shared scratch holds
the previous call's contents), over every candidate filter, with inputs
large enough to need several scratch-sized chunks."""
import ctypes as C

import numpy as np
import pytest

import ferrolearn_amd as F
from ferrolearn_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def blobs(n, d, k, seed, dtype):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-10, 10, size=(k, d))
    X = cen[rng.integers(0, k, size=n)] + rng.normal(size=(n, d))
    return X.astype(dtype), cen


def predict_host(C_, X):
    out = np.zeros(X.shape[0], np.int64)
    dt = L.FL_F32 if X.dtype == np.float32 else L.FL_F64
    st = L.lib().fl_kmeans_predict(C_.ctypes.data_as(C.c_void_p), C_.shape[0], C_.shape[1],
                                   X.ctypes.data_as(C.c_void_p), dt, X.shape[0], X.shape[1],
                                   out.ctypes.data_as(C.c_void_p))
    assert st == L.FL_OK, L.last_error()
    return out


def predict_noalloc(C_, X, precision):
    dC, dX = F.to_device(C_), F.to_device(X)
    out = F.device.empty((X.shape[0],), np.int64)
    dt = L.FL_F32 if X.dtype == np.float32 else L.FL_F64
    st = L.lib().fl_kmeans_predict_device_noalloc(
        C.c_void_p(dC.data_ptr()), C_.shape[0], C_.shape[1], C.c_void_p(dX.data_ptr()), dt,
        X.shape[0], X.shape[1], L.PRECISIONS[precision], C.c_void_p(out.data_ptr()), None)
    assert st == L.FL_OK, L.last_error()
    return out.numpy()


# (n, d, k, dtype, precision): the first call is the largest (BF16x3 fused
# filter over three scratch-sized chunks), then smaller calls over the f64
# filter, the direct exact scan and the single-product BF16 filter.
CALLS = [
    (420_000, 48, 200, np.float32, "bf16x3"),
    (70_000, 33, 100, np.float64, "fp64"),
    (20_000, 16, 20, np.float32, "exact"),
    (9_000, 130, 60, np.float32, "bf16"),
    (3_000, 7, 300, np.float64, "bf16x3"),
]


def test_noalloc_predict_matches_predict():
    for i, (n, d, k, dt, prec) in enumerate(CALLS):
        print(f"call {i}: n={n} d={d} k={k} {np.dtype(dt).name} {prec}", flush=True)
        X, cen = blobs(n, d, k, seed=i, dtype=dt)
        # perturbed true centres plus exact copies of some rows: ties and
        # near-ties between centres
        rng = np.random.default_rng(100 + i)
        C_ = cen + rng.normal(scale=0.3, size=cen.shape)
        C_[-1] = C_[0]
        C_[1] = X[5].astype(np.float64)
        C_ = np.ascontiguousarray(C_)
        got = predict_noalloc(C_, X, prec)
        want = predict_host(C_, X)
        assert (got == want).all(), f"call {i}"
        if n <= 100_000:
            assert (got == O.assign(X.astype(np.float64), C_)[0]).all(), f"call {i}"


def test_noalloc_rejects_oversized_and_empty():
    lib = L.lib()
    # k * d beyond the fixed scratch: InvalidParameter before any launch,
    # the output untouched
    k, d = 8192, 2048
    out = F.to_device(np.full(4, -7, np.int64))
    st = lib.fl_kmeans_predict_device_noalloc(C.c_void_p(1), k, d, C.c_void_p(1), L.FL_F64, 4,
                                              d, L.FL_PREC_FP64, C.c_void_p(out.data_ptr()), None)
    assert st == L.FL_INVALID_PARAMETER
    assert "scratch" in L.last_error()
    assert (out.numpy() == -7).all()
    # feature mismatch, as fl_kmeans_predict reports it
    st = lib.fl_kmeans_predict_device_noalloc(C.c_void_p(1), 4, 3, C.c_void_p(1), L.FL_F64, 4,
                                              5, L.FL_PREC_FP64, C.c_void_p(out.data_ptr()), None)
    assert st == L.FL_SHAPE_MISMATCH
    # no rows: nothing to do
    st = lib.fl_kmeans_predict_device_noalloc(None, 4, 3, None, L.FL_F64, 0, 3, L.FL_PREC_FP64,
                                              None, None)
    assert st == L.FL_OK
