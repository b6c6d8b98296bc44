"""MiniBatchKMeans seed partitions on the GPU (scarf ``ann.py:328-346``).

scarf fits scikit-learn's ``MiniBatchKMeans(n_clusters, random_state, batch_size,
n_init=3)`` with one ``partial_fit`` per block of cells and labels every cell with
``predict``; ``graph_datastore.py:958-975`` stores ``cluster_centers_`` and the labels
under ``kmeans__{n_centroids}__{rand_state}``.  This class has the same
``partial_fit`` / ``predict`` / ``cluster_centers_`` behaviour (scikit-learn 1.7.2,
``sklearn/cluster/_kmeans.py:1678-2303``; dense rows, unit sample weights, k-means++):

* the random state lives on the host and is consumed in scikit-learn's order
  (``randint`` for the init subsample, ``choice`` and ``uniform`` in k-means++,
  ``choice(replace=False)`` for reassignments), so a seed gives scikit-learn's sequence;
* seeding, nearest-center labels and the minibatch center update are HIP kernels
  (``sh_kmeans_pp``, ``sh_kmeans_assign``, ``sh_kmeans_update``) in FP64;
* the per-step reassignment decision (``_kmeans.py:1642-1673``, ``2029-2043``) reads the
  n_clusters weight sums on the host: one small copy per ``partial_fit``.

Rows may be FP32 (promoted exactly) or FP64, on the device or the host; the result is
scikit-learn's on the same values in float64.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _lib
from .device import ptr, require_gpu, stream_handle, workspace

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

__all__ = ["MiniBatchKMeans"]


def _rows(X, device):
    """(2-D device tensor with unit column stride, x_dtype code, leading dimension)."""
    if isinstance(X, torch.Tensor):
        t = X if X.device.type == "cuda" else X.to(device)
        if t.dtype not in (torch.float32, torch.float64):
            t = t.to(torch.float64)
    else:
        a = np.asarray(X)
        if a.dtype != np.float32:
            a = a.astype(np.float64, copy=False)
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device)
    if t.dim() != 2:
        raise ValueError(f"Expected 2D array, got {t.dim()}D array instead")
    if t.shape[1] > 0 and t.stride(1) != 1:
        t = t.contiguous()
    code = _lib.SH_F32 if t.dtype == torch.float32 else _lib.SH_F64
    ld = t.stride(0) if t.shape[0] > 1 else max(t.shape[1], 1)
    return t, code, ld


class MiniBatchKMeans:
    """scikit-learn's ``MiniBatchKMeans`` (``partial_fit`` / ``predict``) on one GPU."""

    def __init__(self, n_clusters: int = 8, *, init: str = "k-means++", max_iter: int = 100,
                 batch_size: int = 1024, verbose: int = 0, compute_labels: bool = True,
                 random_state=None, tol: float = 0.0, max_no_improvement: int = 10,
                 init_size: Optional[int] = None, n_init="auto",
                 reassignment_ratio: float = 0.01, device: str = "cuda"):
        if init != "k-means++":
            raise NotImplementedError("only init='k-means++' (scarf's setting) is supported")
        if reassignment_ratio < 0:
            raise ValueError(f"reassignment_ratio should be >= 0, got {reassignment_ratio} "
                             "instead.")
        self.n_clusters = int(n_clusters)
        self.init = init
        self.max_iter = max_iter
        self.batch_size = int(batch_size)
        self.verbose = verbose
        self.compute_labels = compute_labels
        self.random_state = random_state
        self.tol = tol
        self.max_no_improvement = max_no_improvement
        self.init_size = init_size
        self.n_init = n_init
        self.reassignment_ratio = float(reassignment_ratio)
        self.device = device
        self.n_steps_ = 0
        self._C = None
        self._centers_host = None
        self._labels_dev = None
        self._bufs = {}

    def _buf(self, name: str, numel: int, dtype, device):
        """A per-instance device buffer of at least numel elements (reused across calls on
        the current stream; kernels on one stream run in order)."""
        b = self._bufs.get(name)
        if b is None or b.numel() < numel:
            b = self._bufs[name] = torch.empty(max(int(numel), 1), dtype=dtype, device=device)
        return b[:numel]

    def _assign(self, t, code, ld, n, d, labels, stream):
        k = self.n_clusters
        cn = self._buf("cn", k, torch.float64, t.device)
        wsb = _lib.lib().sh_kmeans_assign_workspace(n, k)
        ws = self._buf("assign_ws", wsb, torch.uint8, t.device) if wsb else None
        _lib.call("sh_kmeans_assign", ptr(t), code, n, d, ld, ptr(self._C), k, ptr(cn),
                  ptr(labels), ptr(ws), wsb, stream)

    # -- state ---------------------------------------------------------------------------
    def _rs(self) -> np.random.RandomState:
        if not hasattr(self, "_random_state"):
            rs = self.random_state
            if rs is None or isinstance(rs, (int, np.integer)):
                rs = np.random.RandomState(rs)
            elif not isinstance(rs, np.random.RandomState):
                raise ValueError(f"{rs!r} cannot be used to seed a numpy.random.RandomState "
                                 "instance")
            self._random_state = rs
        return self._random_state

    @property
    def cluster_centers_(self) -> np.ndarray:
        if self._C is None:
            raise AttributeError("cluster_centers_ (call partial_fit first)")
        if self._centers_host is None:
            self._centers_host = self._C.cpu().numpy()
        return self._centers_host

    @property
    def cluster_centers_tensor(self) -> "torch.Tensor":
        return self._C

    @property
    def labels_(self) -> np.ndarray:
        if self._labels_dev is None:
            raise AttributeError("labels_ (compute_labels=False or no partial_fit yet)")
        return self._labels_dev.cpu().numpy()

    # -- seeding (_kmeans.py:1003-1020 and 174-272) -----------------------------------------
    def _init_centers(self, t, code, ld, rs, stream):
        n, d = t.shape
        k = self.n_clusters
        self._batch_size = min(self.batch_size, n)
        init_size = self.init_size
        if init_size is None:
            init_size = 3 * self._batch_size
            if init_size < k:
                init_size = 3 * k
        elif init_size < k:
            init_size = 3 * k
        init_size = min(init_size, n)
        sub = None
        m = n
        if init_size < n:
            sub = rs.randint(0, n, init_size)
            m = init_size
        trials = 2 + int(np.log(k))
        first = int(rs.choice(m, p=np.ones(m) / m))
        U = rs.uniform(size=(k - 1) * trials) if k > 1 else np.zeros(1)
        dev = t.device
        U_d = torch.from_numpy(np.ascontiguousarray(U, dtype=np.float64)).to(dev)
        rows_d = None if sub is None else torch.from_numpy(sub.astype(np.int64)).to(dev)
        idx_d = torch.empty(k, dtype=torch.int64, device=dev)
        ws, nbytes = workspace("kmeans").get(_lib.lib().sh_kmeans_pp_workspace(m, trials))
        _lib.call("sh_kmeans_pp", ptr(t), code, m, d, ld, ptr(rows_d), k, trials, first,
                  ptr(U_d), ptr(idx_d), ptr(ws), nbytes, stream)
        idx = idx_d.cpu().numpy()
        src = idx if sub is None else sub[idx]
        src_d = torch.from_numpy(np.ascontiguousarray(src, dtype=np.int64)).to(dev)
        self._C = torch.empty((k, d), dtype=torch.float64, device=dev)
        _lib.call("sh_kmeans_gather", ptr(t), code, d, ld, ptr(src_d), None, k, ptr(self._C),
                  stream)
        self.init_indices_ = src
        self._ws_d = torch.zeros(k, dtype=torch.float64, device=dev)
        self._counts = np.zeros(k)
        self._n_since_last_reassign = 0
        self._n_features = d

    # -- partial_fit (_kmeans.py:2203-2303) ----------------------------------------------
    def partial_fit(self, X, y=None, sample_weight=None):
        require_gpu()
        if sample_weight is not None:
            raise NotImplementedError("unit sample weights only (scarf passes none)")
        t, code, ld = _rows(X, self.device)
        n, d = t.shape
        k = self.n_clusters
        rs = self._rs()
        stream = stream_handle()
        if self._C is None:
            if n < k:
                raise ValueError(f"n_samples={n} should be >= n_clusters={k}.")
            self._init_centers(t, code, ld, rs, stream)
        elif d != self._n_features:
            raise ValueError(f"X has {d} features, but MiniBatchKMeans is expecting "
                             f"{self._n_features} features as input.")
        # _random_reassign() is evaluated before the step (_kmeans.py:2029-2043)
        self._n_since_last_reassign += self._batch_size
        reassign = False
        if (self._counts == 0).any() or self._n_since_last_reassign >= 10 * k:
            self._n_since_last_reassign = 0
            reassign = True
        dev = t.device
        labels = self._buf("labels", n, torch.int32, dev)
        counts = self._buf("counts", k, torch.int32, dev)
        self._assign(t, code, ld, n, d, labels, stream)
        _lib.call("sh_kmeans_update", ptr(t), code, n, d, ld, ptr(labels), ptr(self._C), k,
                  ptr(self._ws_d), ptr(counts), stream)
        ws = self._ws_d.cpu().numpy().copy()
        if reassign and self.reassignment_ratio > 0:
            to_reassign = ws < self.reassignment_ratio * ws.max()
            if to_reassign.sum() > 0.5 * n:
                keep = np.argsort(ws)[int(0.5 * n):]
                to_reassign[keep] = False
            n_re = int(to_reassign.sum())
            if n_re:
                new = rs.choice(n, replace=False, size=n_re)
                if self.verbose:
                    print(f"[MiniBatchKMeans] Reassigning {n_re} cluster centers.")
                src_d = torch.from_numpy(np.ascontiguousarray(new, dtype=np.int64)).to(dev)
                dst_d = torch.from_numpy(np.flatnonzero(to_reassign).astype(np.int32)).to(dev)
                _lib.call("sh_kmeans_gather", ptr(t), code, d, ld, ptr(src_d), ptr(dst_d),
                          n_re, ptr(self._C), stream)
            ws[to_reassign] = np.min(ws[~to_reassign])
            self._ws_d.copy_(torch.from_numpy(ws))
        self._counts = ws
        self._centers_host = None
        if self.compute_labels:
            self._labels_dev = self.predict_tensor(t)
        self.n_steps_ += 1
        return self

    # -- predict (_kmeans.py:1066-1098) -------------------------------------------------
    def predict_tensor(self, X) -> "torch.Tensor":
        """int32 labels on the device."""
        require_gpu()
        if self._C is None:
            raise RuntimeError("This MiniBatchKMeans instance is not fitted yet.")
        t, code, ld = _rows(X, self.device)
        n, d = t.shape
        if d != self._n_features:
            raise ValueError(f"X has {d} features, but MiniBatchKMeans is expecting "
                             f"{self._n_features} features as input.")
        labels = torch.empty(n, dtype=torch.int32, device=t.device)
        self._assign(t, code, ld, n, d, labels, stream_handle())
        return labels

    def predict(self, X) -> np.ndarray:
        return self.predict_tensor(X).cpu().numpy()
