"""KMeans with the reference's builder API, running on the gfx950 library.

Mirrors ferrolearn-cluster/src/kmeans.rs: `KMeans(n_clusters)` with
`with_max_iter / with_tol / with_n_init / with_random_state` (kmeans.rs:86-112),
`fit(x)` -> `FittedKMeans` (Fit, kmeans.rs:403-513) exposing
`cluster_centers() / labels() / inertia() / n_iter()` (:121-156), `predict`
(:528-553), `transform` (:555-618) and `KMeans.fit_predict` (:522-525).

Inputs may be numpy arrays (copied to the GPU) or CUDA tensors already in
HBM (float32 or float64, C-contiguous, row-major).
"""
from __future__ import annotations

import copy
import ctypes as C

import numpy as np

from . import _lib as L
from .errors import check


def _dtype_code(dt) -> int:
    name = str(dt)
    if name.endswith("float32"):
        return L.FL_F32
    if name.endswith("float64"):
        return L.FL_F64
    raise TypeError(f"unsupported dtype {dt}: use float32 or float64")


def _host2d(x) -> np.ndarray:
    x = np.asarray(x)
    if x.dtype not in (np.float32, np.float64):
        x = x.astype(np.float64)
    if x.ndim != 2:
        raise ValueError("expected a 2-D array")
    return np.ascontiguousarray(x)


def _is_device(x) -> bool:
    return hasattr(x, "is_cuda") and bool(x.is_cuda)


def _dev2d(x):
    """A device input as the kernels read it: 2-D, row-major and packed.

    The C entry points take a bare pointer and read n*d consecutive values,
    so a strided or broadcast CUDA tensor view (a column slice, a transpose,
    x[::2], an expand) is first copied into a packed tensor; reading the view
    in place would give wrong rows, or read past its storage.  DeviceArray
    is always packed."""
    from .device import DeviceArray
    if len(x.shape) != 2:
        from .errors import ShapeMismatch
        raise ShapeMismatch(f"expected a 2-D array, got shape {tuple(x.shape)}")
    if isinstance(x, DeviceArray):
        return x
    if hasattr(x, "is_contiguous") and not x.is_contiguous():
        x = x.contiguous()
    return x


def _to_host(x) -> np.ndarray:
    """Copy a DeviceArray or CUDA tensor back to a numpy array."""
    from .device import DeviceArray
    if isinstance(x, DeviceArray):
        return x.numpy()
    return x.cpu().numpy()


def _dev_empty(shape, np_dtype, like):
    """Output buffer in HBM of the same kind as the input `like`."""
    from .device import DeviceArray
    if isinstance(like, DeviceArray):
        return DeviceArray(shape, np_dtype)
    import torch
    tdt = {np.dtype(np.float64): torch.float64, np.dtype(np.int64): torch.int64,
           np.dtype(np.float32): torch.float32}[np.dtype(np_dtype)]
    return torch.empty(shape, dtype=tdt, device=like.device)


def _dev_from_host(a, like):
    from .device import DeviceArray, to_device
    if isinstance(like, DeviceArray):
        return to_device(a)
    import torch
    return torch.as_tensor(a, device=like.device).contiguous()


class KMeans:
    """kmeans.rs:55-113.  Defaults from KMeans::new: max_iter 300, tol 1e-4,
    n_init 10, random_state None.  `precision` picks the candidate filter
    ('fp64' GEMM form, 'bf16x3' / 'bf16' BF16 MFMA filters, 'exact' direct folds);
    labels are exact either way."""

    def __init__(self, n_clusters: int):
        self.n_clusters = int(n_clusters)
        self.max_iter = 300
        self.tol = 1e-4
        self.n_init = 10
        self.random_state = None
        self.precision = "fp64"
        self.rng = "library"

    def _with(self, **kw):
        o = copy.copy(self)
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def with_max_iter(self, max_iter: int):
        return self._with(max_iter=int(max_iter))

    def with_tol(self, tol: float):
        return self._with(tol=float(tol))

    def with_n_init(self, n_init: int):
        return self._with(n_init=int(n_init))

    def with_random_state(self, seed: int):
        return self._with(random_state=int(seed))

    def with_precision(self, precision: str):
        return self._with(precision=precision)

    def with_rng(self, rng: str):
        """'library' (default): the library's generator with parallel
        two-level sums in k-means++; 'stdrng': the reference's StdRng stream
        (ChaCha12, rand 0.9) with its plain left folds, so that a seeded fit
        draws the reference's initial centres."""
        if rng not in L.RNGS:
            raise ValueError(f"unknown rng {rng!r}: use 'library' or 'stdrng'")
        return self._with(rng=rng)

    def _params(self):
        p = L.fl_kmeans_params()
        L.lib().fl_kmeans_params_default(C.byref(p), self.n_clusters)
        p.max_iter = self.max_iter
        p.tol = self.tol
        p.n_init = self.n_init
        if self.random_state is not None:
            p.random_state = self.random_state
            p.has_random_state = 1
        p.precision = L.PRECISIONS[self.precision]
        p.rng = L.RNGS[self.rng]
        return p

    def fit(self, x, y=None, init_centers=None) -> "FittedKMeans":
        """Fit<Array2<F>, ()>.  init_centers: optional (n_init, k, d) or
        (k, d) array of initial centres (otherwise the library's k-means++)."""
        return self._fit(x, init_centers, L.lib().fl_kmeans_fit, L.lib().fl_kmeans_fit_device)

    def _fit(self, x, init_centers, fit_host, fit_dev):
        p = self._params()
        rep = L.fl_report()
        inertia = C.c_double()
        n_iter = C.c_int64()
        if _is_device(x):
            x = _dev2d(x)
            n, d = x.shape
            dt = _dtype_code(x.dtype)
            k = self.n_clusters
            centers = _dev_empty((max(k, 0), d), np.float64, x)
            labels = _dev_empty((n,), np.int64, x)
            init_t = None
            if init_centers is not None:
                init_t = _dev_from_host(np.ascontiguousarray(init_centers, dtype=np.float64), x)
            st = fit_dev(
                C.c_void_p(x.data_ptr()), dt, n, d, C.byref(p),
                C.c_void_p(init_t.data_ptr()) if init_t is not None else None,
                C.c_void_p(centers.data_ptr()), C.c_void_p(labels.data_ptr()),
                C.byref(inertia), C.byref(n_iter), C.byref(rep))
            check(st)
            return FittedKMeans(centers, labels, inertia.value, n_iter.value, rep.as_dict(),
                                self.precision)
        xh = _host2d(x)
        n, d = xh.shape
        k = self.n_clusters
        init = None
        if init_centers is not None:
            init = np.ascontiguousarray(np.asarray(init_centers, dtype=np.float64))
            if init.size != self.n_init * k * d:
                raise ValueError("init_centers must hold n_init * n_clusters * n_features values")
        centers = np.zeros((max(k, 0), d), np.float64)
        labels = np.zeros(n, np.int64)
        st = fit_host(
            xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype), n, d, C.byref(p),
            init.ctypes.data_as(C.c_void_p) if init is not None else None,
            centers.ctypes.data_as(C.c_void_p), labels.ctypes.data_as(C.c_void_p),
            C.byref(inertia), C.byref(n_iter), C.byref(rep))
        check(st)
        return FittedKMeans(centers, labels, inertia.value, n_iter.value, rep.as_dict(),
                            self.precision)

    def fit_predict(self, x):
        """kmeans.rs:522-525: fit, then predict against the final centres."""
        return self.fit(x).predict(x)


class FittedKMeans:
    """kmeans.rs:121-156."""

    def __init__(self, centers, labels, inertia, n_iter, report, precision="fp64"):
        self.cluster_centers_ = centers
        self.labels_ = labels
        self.inertia_ = float(inertia)
        self.n_iter_ = int(n_iter)
        self.report = report
        self.precision = precision

    def cluster_centers(self):
        return self.cluster_centers_

    def labels(self):
        return self.labels_

    def inertia(self) -> float:
        return self.inertia_

    def n_iter(self) -> int:
        return self.n_iter_

    def _centers_host(self):
        c = self.cluster_centers_
        if _is_device(c):
            c = _to_host(c)
        return np.ascontiguousarray(c, dtype=np.float64)

    def predict(self, x):
        """Predict (kmeans.rs:528-553): ShapeMismatch on the feature count."""
        c = self._centers_host()
        xh = _host2d(_to_host(x) if _is_device(x) else x)
        n, d = xh.shape
        out = np.zeros(n, np.int64)
        check(L.lib().fl_kmeans_predict(c.ctypes.data_as(C.c_void_p), c.shape[0], c.shape[1],
                                        xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype),
                                        n, d, out.ctypes.data_as(C.c_void_p)))
        return out

    def transform(self, x):
        """Transform (kmeans.rs:555-618): n x k Euclidean distances."""
        c = self._centers_host()
        xh = _host2d(_to_host(x) if _is_device(x) else x)
        n, d = xh.shape
        out = np.zeros((n, c.shape[0]), np.float64)
        check(L.lib().fl_kmeans_transform(c.ctypes.data_as(C.c_void_p), c.shape[0], c.shape[1],
                                          xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype),
                                          n, d, out.ctypes.data_as(C.c_void_p)))
        return out


class MiniBatchKMeans(KMeans):
    """mini_batch_kmeans.rs:77-161.  Defaults from MiniBatchKMeans::new
    (:108-118): batch_size 1024, max_iter 100, tol 0, n_init 3, no seed.
    `precision` picks the filter of the final full assignment."""

    def __init__(self, n_clusters: int):
        super().__init__(n_clusters)
        self.batch_size = 1024
        self.max_iter = 100
        self.tol = 0.0
        self.n_init = 3

    def with_batch_size(self, batch_size: int):
        return self._with(batch_size=int(batch_size))

    def _params(self):
        p = L.fl_minibatch_params()
        L.lib().fl_minibatch_params_default(C.byref(p), self.n_clusters)
        p.batch_size = self.batch_size
        p.max_iter = self.max_iter
        p.tol = self.tol
        p.n_init = self.n_init
        if self.random_state is not None:
            p.random_state = self.random_state
            p.has_random_state = 1
        p.precision = L.PRECISIONS[self.precision]
        p.rng = L.RNGS[self.rng]
        return p

    def fit(self, x, y=None, init_centers=None) -> "FittedKMeans":
        """Fit<Array2<F>, ()> (mini_batch_kmeans.rs:478-592).  init_centers:
        optional (n_init, k, d) or (k, d) initial centres, otherwise the
        library's k-means++ from each run's generator."""
        return self._fit(x, init_centers, L.lib().fl_minibatch_kmeans_fit,
                         L.lib().fl_minibatch_kmeans_fit_device)


def kmeans_plusplus(x, n_clusters: int, seed: int) -> np.ndarray:
    """The library's k-means++ (kmeans.rs:169-253 with its own generator)."""
    xh = _host2d(x)
    n, d = xh.shape
    out = np.zeros((n_clusters, d), np.float64)
    check(L.lib().fl_kmeans_plusplus(xh.ctypes.data_as(C.c_void_p), _dtype_code(xh.dtype), n, d,
                                     n_clusters, C.c_uint64(seed),
                                     out.ctypes.data_as(C.c_void_p)))
    return out


def kmeans_fit_batched(problems, n_workers: int = 8):
    """Fit many independent KMeans problems in one call (fl_kmeans_fit_batched).

    `problems` is a sequence of (x, model, init_centers) with x a device array
    (DeviceArray or CUDA tensor), model a configured `KMeans`, and
    init_centers None (the library's k-means++) or n_init*k*d values.  The
    problems run concurrently, each on its own stream; each result equals
    `model.fit(x, init_centers)` bit for bit.  Returns one `FittedKMeans` per
    problem and the call's report.
    """
    from .device import DeviceArray
    arr = (L.fl_kmeans_problem * max(1, len(problems)))()
    keep = []
    outs = []
    for i, (x, model, init) in enumerate(problems):
        if not _is_device(x):
            raise TypeError("kmeans_fit_batched takes device arrays")
        x = _dev2d(x)
        n, d = x.shape
        k = model.n_clusters
        centers = _dev_empty((max(k, 0), d), np.float64, x)
        labels = _dev_empty((n,), np.int64, x)
        it = None
        if init is not None:
            it = _dev_from_host(np.ascontiguousarray(init, dtype=np.float64), x)
        keep.append((centers, labels, it, x))   # x: a packed copy must outlive the call
        pr = arr[i]
        pr.X = x.data_ptr()
        pr.dtype = _dtype_code(x.dtype)
        pr.n, pr.d = n, d
        pr.params = model._params()
        pr.init_centers = it.data_ptr() if it is not None else None
        pr.centers_out = centers.data_ptr()
        pr.labels_out = labels.data_ptr()
        outs.append(model)
    rep = L.fl_report()
    check(L.lib().fl_kmeans_fit_batched(arr, len(problems), int(n_workers), C.byref(rep)))
    fitted = [FittedKMeans(keep[i][0], keep[i][1], arr[i].inertia, arr[i].n_iter, {},
                           outs[i].precision) for i in range(len(problems))]
    return fitted, rep.as_dict()
