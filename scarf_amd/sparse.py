"""The cell graph as a scipy sparse matrix, built on the GPU.

scarf's ``load_graph`` (``graph_datastore.py:1022-1075``) reads the stored ``edges`` and
``weights`` through ``_store_to_sparse`` (``graph_datastore.py:474-511``): the first
``use_k`` neighbours of every cell, a CSR matrix, optionally symmetrised as
``g + g.T - g.multiply(g.T)`` and cut to its upper triangle (``scipy.sparse.triu``, which
returns COO).  ``sh_graph_csr`` does the sorting, duplicate summing, transpose and
symmetrisation on the device; the result is copied back as scipy's arrays (int32 indices,
float64 data, sorted columns).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import ptr, require_gpu, stream_handle, workspace

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

__all__ = ["graph_csr"]


def graph_csr(edges, weights, n_cells: int, k: int, use_k=None, symmetric: bool = False,
              upper_only: bool = False):
    """scipy matrix of the kNN graph: CSR, or COO when ``symmetric and upper_only``."""
    require_gpu()
    from scipy.sparse import coo_matrix, csr_matrix

    n, k = int(n_cells), int(k)
    if use_k is None or use_k > k:
        use_k = k
    if use_k < 1:
        use_k = 1
    e = np.asarray(edges)
    if e.ndim != 2 or e.shape[1] != 2:
        raise ValueError("edges must be an (n_edges, 2) array of (row, col)")
    if e.dtype.kind not in "iu":
        raise ValueError("edges must hold integers")
    w = np.asarray(weights, dtype=np.float64).ravel()
    nnz = e.shape[0]
    if w.shape[0] != nnz or nnz % k:
        raise ValueError("edges and weights must hold n_cells * k entries")
    # uint64 (scarf's "u8" edges) and int64 share bits for cell indices: view, do not copy
    e64 = np.ascontiguousarray(e.view(np.int64) if e.dtype == np.uint64 else
                               e.astype(np.int64, copy=False))
    dev = torch.device("cuda")
    e_d = torch.from_numpy(e64).to(dev)
    w_d = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    m = nnz // k * use_k
    cap = max(2 * m if symmetric else m, 1)
    indptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
    indices = torch.empty(cap, dtype=torch.int32, device=dev)
    data = torch.empty(cap, dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    L = _lib.lib()
    wsb = L.sh_graph_csr_workspace(n, nnz, k, use_k)
    ws, nbytes = workspace("graph").get(wsb)
    _lib.call("sh_graph_csr", ptr(e_d), ptr(w_d), nnz, k, use_k, n, int(bool(symmetric)),
              int(bool(upper_only)), ptr(indptr), ptr(indices), ptr(data), cap, ptr(bad),
              ptr(ws), nbytes, stream_handle())
    ip = indptr.cpu().numpy()
    if int(bad.item()):
        raise ValueError(f"{int(bad.item())} edges name a cell outside [0, {n})")
    total = int(ip[-1])
    idx_t = np.int32 if max(n, total) < 2 ** 31 else np.int64
    g = csr_matrix((data[:total].cpu().numpy(),
                    indices[:total].cpu().numpy().astype(idx_t, copy=False),
                    ip.astype(idx_t)), shape=(n, n), copy=False)
    g.has_sorted_indices = True
    if symmetric and upper_only:
        return coo_matrix(g)
    return g
