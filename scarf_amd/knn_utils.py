"""Drop-ins for scarf's ``knn_utils.self_query_knn`` and ``smoothen_dists``.

``store`` is any writable mapping (a zarr group in scarf; a dict works); the
datasets written and their dtypes follow the reference: ``indices`` u8 and
``distances`` f8 of shape (n, k) (``knn_utils.py:54-72``), ``edges`` (u8, u8) of shape
(n*k, 2) and ``weights`` f8 of shape (n*k,) (``knn_utils.py:108-158``).
"""
from __future__ import annotations

import numpy as np

from .ann import AnnStream
from .smooth import edges_of, smooth_knn

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

__all__ = ["self_query_knn", "smoothen_dists"]


def self_query_knn(ann_obj: AnnStream, store, chunk_size: int, nthreads: int = 1) -> float:
    """k+1 self query of every cell, self repair, recall % (``knn_utils.py:23-76``).

    All cells are queried in one pass: with the exact re-rank, each row's result does
    not depend on how the rows are batched.
    """
    n = ann_obj.nCells
    k = ann_obj.k
    E = ann_obj.annIdx.embedding()
    idx, dist, flags = ann_obj.annIdx.knn_query_tensor(E, k, self_offset=0)
    n_mis = int(((flags & 4) > 0).sum().item())
    store["indices"] = idx.cpu().numpy().astype(np.uint64)
    store["distances"] = dist.cpu().numpy().astype(np.float64)
    return 100.0 * (n - n_mis) / n


def smoothen_dists(store, z_idx, z_dist, lc: float, bw: float, chunk_size: int):
    """Edge list and fuzzy weights of the kNN graph (``knn_utils.py:89-159``)."""
    if isinstance(z_idx, torch.Tensor):
        ki = z_idx.to(torch.int64)
    else:
        ki = torch.from_numpy(np.asarray(z_idx).astype(np.int64)).cuda()
    if isinstance(z_dist, torch.Tensor):
        kd = z_dist.to(torch.float32)
    else:
        kd = torch.from_numpy(np.asarray(z_dist, dtype=np.float32)).cuda()
    n, k = ki.shape
    _, _, w = smooth_knn(kd, ki, chunk_size, lc, bw)
    store["edges"] = edges_of(n, k, ki.cpu().numpy())
    store["weights"] = w.cpu().numpy().astype(np.float64)
    return None
