"""euclidean_distances, pairwise_distances_argmin and
pairwise_distances_argmin_min on the GPU.

Mirrors ferrolearn-metrics/src/pairwise.rs:157-189 and :564-629 with
Metric::Euclidean: the distances are euclidean_distances' expansion,
sqrt(max((|x|^2 + |y|^2) - 2 x.y, 0)) with left folds, and argmin ties go to
the lowest index.  Results are bit-exact against the reference's f64 path.
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np

from . import _lib as L
from .cluster import _dev2d, _dev_empty, _dtype_code, _host2d, _is_device
from .errors import InvalidParameter, check


class Metric(enum.Enum):
    """pairwise.rs:22-31.  Only Euclidean is on the squared-Euclidean path."""
    Euclidean = 0
    Manhattan = 1
    Cosine = 2
    Chebyshev = 3


def _run(x, y, metric, want_min):
    if metric is not Metric.Euclidean:
        raise InvalidParameter(f"InvalidParameter: metric: {metric.name} is not on the "
                               "squared-Euclidean path; only Metric.Euclidean runs here")
    if _is_device(x) or _is_device(y):
        import torch
        from .device import DeviceArray
        xd = _dev2d(x if _is_device(x) else torch.from_numpy(_host2d(x)).cuda())
        yd = _dev2d(y if _is_device(y) else torch.from_numpy(_host2d(y)).cuda())
        if str(xd.dtype) != str(yd.dtype):   # f32 widened to f64 is exact
            if isinstance(xd, DeviceArray) or isinstance(yd, DeviceArray):
                raise TypeError("X and Y must share a dtype")
            xd, yd = xd.double(), yd.double()
        n, dx = xd.shape
        m, dy = yd.shape
        idx = _dev_empty((n,), np.int64, xd)
        mins = _dev_empty((n,), np.float64, xd) if want_min else None
        check(L.lib().fl_pairwise_distances_argmin_min_device(
            C.c_void_p(xd.data_ptr()), _dtype_code(xd.dtype), n, dx, C.c_void_p(yd.data_ptr()),
            m, dy, C.c_void_p(idx.data_ptr()),
            C.c_void_p(mins.data_ptr()) if want_min else None))
        return idx, mins
    xh, yh = _host2d(x), _host2d(y)
    if xh.dtype != yh.dtype:
        xh, yh = xh.astype(np.float64), yh.astype(np.float64)
    n, dx = xh.shape
    m, dy = yh.shape
    idx = np.zeros(n, np.int64)
    mins = np.zeros(n, np.float64) if want_min else None
    check(L.lib().fl_pairwise_distances_argmin_min(
        xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype), n, dx,
        yh.ctypes.data_as(C.c_void_p), m, dy, idx.ctypes.data_as(C.c_void_p),
        mins.ctypes.data_as(C.c_void_p) if want_min else None))
    return idx, mins


def _pair(x, y):
    """The two inputs on one side (device when either is) with one dtype."""
    if _is_device(x) or _is_device(y):
        import torch
        from .device import DeviceArray
        xd = _dev2d(x if _is_device(x) else torch.from_numpy(_host2d(x)).cuda())
        yd = _dev2d(y if _is_device(y) else torch.from_numpy(_host2d(y)).cuda())
        if str(xd.dtype) != str(yd.dtype):   # f32 widened to f64 is exact
            if isinstance(xd, DeviceArray) or isinstance(yd, DeviceArray):
                raise TypeError("X and Y must share a dtype")
            xd, yd = xd.double(), yd.double()
        return True, xd, yd
    xh, yh = _host2d(x), _host2d(y)
    if xh.dtype != yh.dtype:
        xh, yh = xh.astype(np.float64), yh.astype(np.float64)
    return False, xh, yh


def euclidean_distances(x, y):
    """pairwise.rs:157-189: the (n, m) f64 matrix of distances between the
    rows of x and y (a device tensor when either input is on the device)."""
    dev, xa, ya = _pair(x, y)
    n, dx = xa.shape
    m, dy = ya.shape
    if dev:
        out = _dev_empty((n, m), np.float64, xa)
        check(L.lib().fl_euclidean_distances_device(
            C.c_void_p(xa.data_ptr()), _dtype_code(xa.dtype), n, dx, C.c_void_p(ya.data_ptr()),
            m, dy, C.c_void_p(out.data_ptr())))
        return out
    out = np.zeros((n, m), np.float64)
    check(L.lib().fl_euclidean_distances(
        xa.ctypes.data_as(C.c_void_p), _dtype_code(xa.dtype), n, dx,
        ya.ctypes.data_as(C.c_void_p), m, dy, out.ctypes.data_as(C.c_void_p)))
    return out


def pairwise_distances_argmin(x, y, metric=Metric.Euclidean):
    """pairwise.rs:564-588: for each row of x, the index of the closest row
    of y."""
    return _run(x, y, metric, False)[0]


def pairwise_distances_argmin_min(x, y, metric=Metric.Euclidean):
    """pairwise.rs:596-629: (indices, minimum distances)."""
    return _run(x, y, metric, True)
