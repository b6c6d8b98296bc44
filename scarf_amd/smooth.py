"""kNN edge weights (scarf ``smoothen_dists``, ``knn_utils.py:89-159``) on the GPU."""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import ptr, require_gpu, stream_handle, workspace

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def smooth_knn(dist: "torch.Tensor", idx: "torch.Tensor", chunk_size: int,
               local_connectivity: float = 1.0, bandwidth: float = 1.5, n_iter: int = 64,
               stream=None):
    """(sigma f32 [n], rho f32 [n], weights f32 [n*k]) with the zero -> global-min rule."""
    require_gpu()
    n, k = dist.shape
    dev = dist.device
    dist = dist.contiguous().to(torch.float32)
    idx = idx.contiguous().to(torch.int64)
    sig = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
    rho = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
    w = torch.empty(max(n * k, 1), dtype=torch.float32, device=dev)
    if n == 0:
        return sig[:0], rho[:0], w[:0]
    buf, nbytes = workspace("smooth").get(_lib.lib().sh_smooth_knn_workspace(n, chunk_size))
    _lib.call("sh_smooth_knn", ptr(dist), ptr(idx), n, k, int(chunk_size),
              float(local_connectivity), float(bandwidth), int(n_iter), ptr(sig), ptr(rho),
              ptr(w), ptr(buf), nbytes, stream_handle(stream))
    return sig[:n], rho[:n], w[:n * k]


def edges_of(n: int, k: int, idx: np.ndarray) -> np.ndarray:
    """scarf's edge list: row-major (i, idx[i, j]) pairs as uint64 [n*k, 2]."""
    e = np.empty((n * k, 2), dtype=np.uint64)
    e[:, 0] = np.repeat(np.arange(n, dtype=np.uint64), k)
    e[:, 1] = np.asarray(idx).reshape(-1).astype(np.uint64)
    return e
