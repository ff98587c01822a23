"""ctypes binding of libferrolearn_hip.so (the C ABI in include/ferrolearn_hip.h).

The library is the product: there is no CPU fallback.  If the shared object
is missing the import fails with an error that says how to build it.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libferrolearn_hip.so")

FL_OK = 0
FL_SHAPE_MISMATCH = 1
FL_INSUFFICIENT_SAMPLES = 2
FL_INVALID_PARAMETER = 3
FL_NUMERICAL_INSTABILITY = 4
FL_DEVICE_ERROR = 5

FL_F32 = 0
FL_F64 = 1

FL_PREC_FP64 = 0
FL_PREC_BF16 = 1
FL_PREC_BF16X3 = 2
FL_PREC_EXACT = 3
FL_PREC_FP8 = 4

FL_UNIFORM = 0
FL_DISTANCE = 1
FL_GRAPH_CONNECTIVITY = 0
FL_RNG_LIBRARY = 0
FL_RNG_STDRNG = 1
RNGS = {"library": FL_RNG_LIBRARY, "stdrng": FL_RNG_STDRNG}
FL_GRAPH_DISTANCE = 1

PRECISIONS = {"fp64": FL_PREC_FP64, "bf16": FL_PREC_BF16, "bf16x3": FL_PREC_BF16X3,
              "exact": FL_PREC_EXACT, "fp8": FL_PREC_FP8}


class fl_report(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double),
        ("ms_distance", C.c_double),
        ("ms_finalize", C.c_double),
        ("ms_fallback", C.c_double),
        ("ms_update", C.c_double),
        ("n_rows", C.c_int64),
        ("n_candidates", C.c_int64),
        ("n_fallback", C.c_int64),
        ("n_iter", C.c_int64),
        ("n_dist_launches", C.c_int64),
        ("n_second", C.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class fl_kmeans_params(C.Structure):
    _fields_ = [
        ("n_clusters", C.c_int64),
        ("max_iter", C.c_int64),
        ("tol", C.c_double),
        ("n_init", C.c_int64),
        ("random_state", C.c_uint64),
        ("has_random_state", C.c_int32),
        ("precision", C.c_int32),
        ("rng", C.c_int32),
        ("reserved", C.c_int32),
    ]


class fl_bisecting_params(C.Structure):
    _fields_ = [
        ("n_clusters", C.c_int64),
        ("max_iter", C.c_int64),
        ("n_init", C.c_int64),
        ("random_state", C.c_uint64),
        ("has_random_state", C.c_int32),
        ("strategy", C.c_int32),
    ]


class fl_minibatch_params(C.Structure):
    _fields_ = [
        ("n_clusters", C.c_int64),
        ("batch_size", C.c_int64),
        ("max_iter", C.c_int64),
        ("tol", C.c_double),
        ("n_init", C.c_int64),
        ("random_state", C.c_uint64),
        ("has_random_state", C.c_int32),
        ("precision", C.c_int32),
        ("rng", C.c_int32),
        ("reserved", C.c_int32),
    ]


class fl_kmeans_problem(C.Structure):
    _fields_ = [
        ("X", C.c_void_p),
        ("dtype", C.c_int32),
        ("n", C.c_int64),
        ("d", C.c_int64),
        ("params", fl_kmeans_params),
        ("init_centers", C.c_void_p),
        ("centers_out", C.c_void_p),
        ("labels_out", C.c_void_p),
        ("inertia", C.c_double),
        ("n_iter", C.c_int64),
        ("status", C.c_int32),
    ]

FL_ROUTE_NETWORK = 0
FL_ROUTE_BOUNDS = 1
FL_ROUTE_CENTRES = 2
ROUTES = {"network": FL_ROUTE_NETWORK, "bounds": FL_ROUTE_BOUNDS, "centres": FL_ROUTE_CENTRES}


class fl_router_params(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int64),
        ("cell_iter", C.c_int64),
        ("hidden", C.c_int64),
        ("seed", C.c_uint64),
    ]


class fl_router_train_params(C.Structure):
    _fields_ = [
        ("n_neighbors", C.c_int64),
        ("epochs", C.c_int64),
        ("batch_size", C.c_int64),
        ("learning_rate", C.c_double),
        ("alpha", C.c_double),
        ("seed", C.c_uint64),
        ("target", C.c_int32),
    ]


FL_TARGET_CERTIFY = 0
FL_TARGET_NEIGHBORS = 1
TARGETS = {"certify": FL_TARGET_CERTIFY, "neighbors": FL_TARGET_NEIGHBORS}


class fl_router_report(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double),
        ("ms_cells", C.c_double),
        ("ms_targets", C.c_double),
        ("ms_train", C.c_double),
        ("ms_route", C.c_double),
        ("ms_bounds", C.c_double),
        ("ms_scan", C.c_double),
        ("ms_merge", C.c_double),
        ("ms_fallback", C.c_double),
        ("n_queries", C.c_int64),
        ("n_rescan", C.c_int64),
        ("n_fallback", C.c_int64),
        ("rows_scanned", C.c_int64),
        ("loss_first", C.c_double),
        ("loss_last", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)
_rp = C.POINTER(fl_report)

# name -> (restype, argtypes); every symbol the header declares.
SIGNATURES = {
    "fl_last_error": (C.c_char_p, []),
    "fl_version": (C.c_int, []),
    "fl_device_count": (C.c_int, []),
    "fl_kmeans_params_default": (None, [C.POINTER(fl_kmeans_params), _i64]),
    "fl_kmeans_fit": (C.c_int, [_vp, _i32, _i64, _i64, C.POINTER(fl_kmeans_params), _vp, _vp,
                                _vp, _dp, _ip, _rp]),
    "fl_kmeans_fit_device": (C.c_int, [_vp, _i32, _i64, _i64, C.POINTER(fl_kmeans_params), _vp,
                                       _vp, _vp, _dp, _ip, _rp]),
    "fl_kmeans_predict": (C.c_int, [_vp, _i64, _i64, _vp, _i32, _i64, _i64, _vp]),
    "fl_kmeans_fit_batched": (C.c_int, [C.POINTER(fl_kmeans_problem), _i64, _i32, _rp]),
    "fl_kmeans_batched_n_variants": (_i32, []),
    "fl_kmeans_batched_variant": (C.c_int, [_i32, C.POINTER(_i32), C.POINTER(_i32),
                                            C.POINTER(_i32)]),
    "fl_kmeans_batched_set_tuning": (C.c_int, [_vp]),
    "fl_kmeans_batched_get_tuning": (C.c_int, [_vp]),
    "fl_kmeans_transform": (C.c_int, [_vp, _i64, _i64, _vp, _i32, _i64, _i64, _vp]),
    "fl_kmeans_predict_device_noalloc": (C.c_int, [_vp, _i64, _i64, _vp, _i32, _i64, _i64, _i32,
                                                   _vp, _vp]),
    "fl_kmeans_plusplus": (C.c_int, [_vp, _i32, _i64, _i64, _i64, C.c_uint64, _vp]),
    "fl_left_fold_device": (C.c_int, [_vp, _i64, _i64, _i64, _vp, _vp]),
    "fl_router_cell_bounds_device": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _i64, _i32, _vp,
                                               _vp]),
    "fl_minibatch_params_default": (None, [C.POINTER(fl_minibatch_params), _i64]),
    "fl_minibatch_kmeans_fit": (C.c_int, [_vp, _i32, _i64, _i64, C.POINTER(fl_minibatch_params),
                                          _vp, _vp, _vp, _dp, _ip, _rp]),
    "fl_minibatch_kmeans_fit_device": (C.c_int, [_vp, _i32, _i64, _i64,
                                                 C.POINTER(fl_minibatch_params), _vp, _vp, _vp,
                                                 _dp, _ip, _rp]),
    "fl_knn_fit_classifier": (C.c_int, [_vp, _i32, _i64, _i64, _vp, _i64, _i64, _i32, _i32,
                                        C.POINTER(_vp)]),
    "fl_knn_fit_regressor": (C.c_int, [_vp, _i32, _i64, _i64, _vp, _i64, _i64, _i32, _i32,
                                       C.POINTER(_vp)]),
    "fl_knn_predict_classify": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _vp, _rp]),
    "fl_knn_predict_regress": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _vp, _rp]),
    "fl_knn_predict_proba": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _vp, _rp]),
    "fl_knn_kneighbors": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _i64, _vp, _vp, _rp]),
    "fl_knn_kneighbors_device": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _i64, _vp, _vp, _rp]),
    "fl_nn_fit": (C.c_int, [_vp, _i32, _i64, _i64, _i64, _i32, C.POINTER(_vp)]),
    "fl_bisecting_params_default": (None, [C.POINTER(fl_bisecting_params), _i64]),
    "fl_bisecting_kmeans_fit": (C.c_int, [_vp, _i32, _i64, _i64, C.POINTER(fl_bisecting_params),
                                          _vp, _vp, C.POINTER(C.c_double),
                                          C.POINTER(C.c_int64)]),
    "fl_bisecting_kmeans_fit_device": (C.c_int, [_vp, _i32, _i64, _i64,
                                                 C.POINTER(fl_bisecting_params), _vp, _vp,
                                                 C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "fl_euclidean_distances": (C.c_int, [_vp, _i32, _i64, _i64, _vp, _i64, _i64, _vp]),
    "fl_euclidean_distances_device": (C.c_int, [_vp, _i32, _i64, _i64, _vp, _i64, _i64, _vp]),
    "fl_pairwise_distances_argmin_min": (C.c_int, [_vp, _i32, _i64, _i64, _vp, _i64, _i64, _vp,
                                                   _vp]),
    "fl_pairwise_distances_argmin_min_device": (C.c_int, [_vp, _i32, _i64, _i64, _vp, _i64, _i64,
                                                          _vp, _vp]),
    "fl_knn_kneighbors_graph": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _i64, _i32, _vp, _vp, _vp,
                                          _rp]),
    "fl_knn_kneighbors_graph_device": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _i64, _i32, _vp, _vp,
                                                 _vp, _rp]),
    "fl_knn_predict_classify_device": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _vp, _rp]),
    "fl_knn_predict_regress_device": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _vp, _rp]),
    "fl_knn_set_precision": (C.c_int, [_vp, _i32]),
    "fl_knn_n_classes": (_i64, [_vp]),
    "fl_knn_classes": (C.c_int, [_vp, _vp]),
    "fl_knn_n_samples_fit": (_i64, [_vp]),
    "fl_knn_free": (None, [_vp]),
    "fl_router_params_default": (None, [C.POINTER(fl_router_params), _i64]),
    "fl_router_train_params_default": (None, [C.POINTER(fl_router_train_params)]),
    "fl_router_build": (C.c_int, [_vp, C.POINTER(fl_router_params), C.POINTER(_vp),
                                  C.POINTER(fl_router_report)]),
    "fl_router_train_device": (C.c_int, [_vp, _vp, _i32, _i64, _i64,
                                         C.POINTER(fl_router_train_params),
                                         C.POINTER(fl_router_report)]),
    "fl_router_kneighbors_device": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _i64, _i64, _i32,
                                              _vp, _vp, C.POINTER(fl_router_report)]),
    "fl_router_n_params": (_i64, [_vp]),
    "fl_router_get_params": (C.c_int, [_vp, _vp]),
    "fl_router_set_params": (C.c_int, [_vp, _vp]),
    "fl_router_cells": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "fl_router_mlp_gradients": (C.c_int, [_i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64,
                                          C.c_double, _dp, _vp]),
    "fl_router_free": (None, [_vp]),
    "fl_backend_gemm_f64": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _i64, _vp]),
    "fl_device_malloc": (C.c_int, [C.POINTER(_vp), C.c_uint64]),
    "fl_device_free": (C.c_int, [_vp]),
    "fl_memcpy_to_device": (C.c_int, [_vp, _vp, C.c_uint64]),
    "fl_memcpy_to_host": (C.c_int, [_vp, _vp, C.c_uint64]),
    "fl_memcpy_device": (C.c_int, [_vp, _vp, C.c_uint64]),
    "fl_device_synchronize": (C.c_int, []),
}

_lib = None


def lib():
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C ferrolearn_amd/csrc` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        _load_tuning(L)
    return _lib


def _load_tuning(L):
    """Apply the committed tile-shape table of the batched KMeans kernel
    (tuning_gfx950.json, written by ferrolearn_amd.tuning.autotune)."""
    import json
    path = os.path.join(HERE, "tuning_gfx950.json")
    if not os.path.exists(path):
        return
    with open(path) as f:
        doc = json.load(f)
    raw = (C.c_int32 * 16)()
    if L.fl_kmeans_batched_get_tuning(raw) != FL_OK:
        return
    dims, ks = (8, 16, 32, 64), (16, 64, 256, 1024)
    for key, v in doc.get("table", {}).items():
        D, K = (int(x) for x in key.split("x"))
        raw[4 * dims.index(D) + ks.index(K)] = int(v)
    if L.fl_kmeans_batched_set_tuning(raw) != FL_OK:
        raise ImportError(f"{path}: invalid tuning table: {L.fl_last_error().decode()}")


def last_error() -> str:
    return lib().fl_last_error().decode()
