"""Restatement of scarf's graph loading (oracle only).

* ``graph_datastore.py:474-511``  _store_to_sparse: the first ``use_k`` of every group of k
  stored edges (``np.tile([True] * use_k + [False] * (k - use_k), n_cells)``), then
  ``csr_matrix((w, (e[:, 0], e[:, 1])), shape=(n_cells, n_cells))``
* ``graph_datastore.py:1022-1075`` load_graph: ``symmetrize(g) = g + g.T - g.multiply(g.T)``
  and ``scipy.sparse.triu`` for ``upper_only``

scipy (the reference's own dependency) does the sparse arithmetic here exactly as in the
reference.  Pinned in ``tests/test_oracle.py`` by a hand-computed three-cell graph (mutual
and one-way edges, an explicit zero, the a + b - ab rule) and by symmetry of the result.
"""
from __future__ import annotations

import numpy as np


def store_to_sparse(edges, weights, n_cells: int, k: int, use_k=None):
    from scipy.sparse import csr_matrix

    if use_k is None:
        use_k = k
    if use_k > k:
        use_k = k
    if use_k < 1:
        use_k = 1
    w, e = np.asarray(weights), np.asarray(edges)
    if use_k != k:
        indexer = np.tile([True] * use_k + [False] * (k - use_k), n_cells)
        w, e = w[indexer], e[indexer]
    return csr_matrix((w, (e[:, 0], e[:, 1])), shape=(n_cells, n_cells))


def load_graph(edges, weights, n_cells: int, k: int, symmetric=None, upper_only=None,
               use_k=None):
    from scipy.sparse import triu

    g = store_to_sparse(edges, weights, n_cells, k, use_k)
    if symmetric is True:
        t = g + g.T
        g = t - g.multiply(g.T)
        if upper_only is True:
            g = triu(g)
    return g
