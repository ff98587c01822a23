"""Cell router for exact kNN: a learned index over a fitted brute-force model.

No reference counterpart; the answer is the reference's brute-force answer
(kdtree.rs:364-391), bit for bit.  `CellRouter(fitted, n_cells)` partitions
the fitted model's rows into cells with a coarse KMeans; `train(queries)`
fits the router network (layers, loss and Adam of
ferrolearn-neural/src/mlp.rs) on the cells of each training query's exact
neighbours; `kneighbors(x, n_probe=...)` scans n_probe cells per query,
certifies the answer with the triangle inequality, and sends every query it
cannot certify to the brute force.  `route="bounds"` chooses the cells by
their lower bounds instead of the network (the same search without the
learned part).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .cluster import _dev2d, _dev_empty, _dtype_code, _host2d, _is_device
from .errors import check


class CellRouter:
    def __init__(self, fitted, n_cells: int, cell_iter: int = 10, hidden: int = 512,
                 seed: int = 0):
        self._fitted = fitted   # the router reads the fitted model's rows
        p = L.fl_router_params()
        L.lib().fl_router_params_default(C.byref(p), int(n_cells))
        p.cell_iter = int(cell_iter)
        p.hidden = int(hidden)
        p.seed = int(seed)
        h = C.c_void_p()
        rep = L.fl_router_report()
        check(L.lib().fl_router_build(fitted._h, C.byref(p), C.byref(h), C.byref(rep)))
        self._h = h
        self._d = fitted._d
        self.n_cells = int(n_cells)
        self.hidden = int(hidden)
        self.report = rep.as_dict()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().fl_router_free(h)
            self._h = None

    def train(self, queries, n_neighbors: int = 10, epochs: int = 5, batch_size: int = 4096,
              learning_rate: float = 1e-3, alpha: float = 0.0, seed: int = 0,
              target: str = "certify") -> dict:
        """Fit the network to the cells each query's exact n_neighbors call for:
        target "certify" -- every cell whose lower bound does not exceed the
        exact k-th distance; "neighbors" -- the cells holding the neighbours."""
        from .device import to_device
        q = _dev2d(queries) if _is_device(queries) else to_device(_host2d(queries))
        nq, d = q.shape
        p = L.fl_router_train_params()
        L.lib().fl_router_train_params_default(C.byref(p))
        p.n_neighbors, p.epochs, p.batch_size = int(n_neighbors), int(epochs), int(batch_size)
        p.learning_rate, p.alpha, p.seed = float(learning_rate), float(alpha), int(seed)
        p.target = L.TARGETS[target]
        rep = L.fl_router_report()
        check(L.lib().fl_router_train_device(self._h, C.c_void_p(q.data_ptr()),
                                             _dtype_code(q.dtype), nq, d, C.byref(p),
                                             C.byref(rep)))
        self.report = rep.as_dict()
        return self.report

    def kneighbors(self, x, n_neighbors=None, n_probe: int = 8, route: str = "network"):
        """(distances, indices), each n_queries x k, equal to the brute force's."""
        from .device import to_device
        dev = _is_device(x)
        q = _dev2d(x) if dev else to_device(_host2d(x))
        nq, d = q.shape
        k = 0 if n_neighbors is None else int(n_neighbors)
        kk = self._fitted.n_neighbors if k == 0 else k
        dist = _dev_empty((nq, max(kk, 0)), np.float64, q)
        idx = _dev_empty((nq, max(kk, 0)), np.int64, q)
        rep = L.fl_router_report()
        check(L.lib().fl_router_kneighbors_device(
            self._h, C.c_void_p(q.data_ptr()), _dtype_code(q.dtype), nq, d, k, int(n_probe),
            L.ROUTES[route], C.c_void_p(dist.data_ptr()), C.c_void_p(idx.data_ptr()),
            C.byref(rep)))
        self.report = rep.as_dict()
        if dev:
            return dist, idx
        return dist.numpy(), idx.numpy()

    def params(self) -> np.ndarray:
        n = int(L.lib().fl_router_n_params(self._h))
        out = np.zeros(n, np.float32)
        check(L.lib().fl_router_get_params(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != int(L.lib().fl_router_n_params(self._h)):
            raise ValueError("parameter vector has the wrong length")
        check(L.lib().fl_router_set_params(self._h, p.ctypes.data_as(C.c_void_p)))

    def cell_bounds(self, x, kb: int, exact: bool = False):
        """(keys, cells), each n_queries x kb: the kb smallest per-cell lower
        bounds on the distance of a cell's rows to each query, ascending,
        computed as the search does (BF16x3 MFMA lower bounds of the folds
        unless exact, or when the data lie outside the BF16 range)."""
        from .device import to_device, empty
        q = _dev2d(x) if _is_device(x) else to_device(_host2d(x))
        nq, d = q.shape
        keys = empty((nq, int(kb)), np.float64)
        ids = empty((nq, int(kb)), np.int32)
        check(L.lib().fl_router_cell_bounds_device(
            self._h, C.c_void_p(q.data_ptr()), _dtype_code(q.dtype), nq, d, int(kb),
            1 if exact else 0, C.c_void_p(keys.data_ptr()), C.c_void_p(ids.data_ptr())))
        return keys.numpy(), ids.numpy()

    def cells(self):
        """(centres n_cells x d, radii, counts, cell of each database row)."""
        n = self._fitted.n_samples_fit()
        cen = np.zeros((self.n_cells, self._d), np.float64)
        rad = np.zeros(self.n_cells, np.float64)
        cnt = np.zeros(self.n_cells, np.int64)
        lab = np.zeros(n, np.int64)
        check(L.lib().fl_router_cells(self._h, cen.ctypes.data_as(C.c_void_p),
                                      rad.ctypes.data_as(C.c_void_p),
                                      cnt.ctypes.data_as(C.c_void_p),
                                      lab.ctypes.data_as(C.c_void_p)))
        return cen, rad, cnt, lab


def split_params(p, d: int, hidden: int, n_cells: int):
    """W1 (d x h), b1, W2 (h x h), b2, W3 (h x n_cells), b3 from a flat vector."""
    shapes = [(d, hidden), (hidden,), (hidden, hidden), (hidden,), (hidden, n_cells), (n_cells,)]
    out, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n]).reshape(s))
        o += n
    return out


def mlp_gradients(params, X, cells, hidden: int, n_cells: int, alpha: float = 0.0):
    """One forward/backward pass of the router network on the GPU: (loss, grads)."""
    X = np.ascontiguousarray(X, np.float32)
    cells = np.ascontiguousarray(cells, np.int32)
    params = np.ascontiguousarray(params, np.float32)
    nb, d = X.shape
    k = cells.shape[1]
    grads = np.zeros_like(params)
    loss = C.c_double()
    check(L.lib().fl_router_mlp_gradients(d, int(hidden), int(n_cells),
                                          params.ctypes.data_as(C.c_void_p),
                                          X.ctypes.data_as(C.c_void_p), nb,
                                          cells.ctypes.data_as(C.c_void_p), k, float(alpha),
                                          C.byref(loss), grads.ctypes.data_as(C.c_void_p)))
    return loss.value, grads
