"""FP64 brute-force kNN and scarf's self-query repair (oracle only).

References (scarf source tree):
* ``ann.py:14-28``     hnswlib index in space 'l2': squared Euclidean distances
* ``ann.py:194-205``   transform_ann: query k (mapping) or k+1 then fix_knn_query (self)
* ``ann.py:31-52``     fix_knn_query: drop column 0 when it is the query itself, else
                       remove the query where it appears, else drop the last column;
                       count the rows that needed repair
* ``knn_utils.py:23-76``  self_query_knn: recall % = 100 * (n - misses) / n

The distance loop itself lives in ``knn_ref.c`` (built by ``oracle/Makefile``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle_knn.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        lib = ctypes.CDLL(path)
        i64, p = ctypes.c_int64, ctypes.c_void_p
        lib.oracle_knn_f64.argtypes = [p, i64, i64, p, i64, i64, ctypes.c_int, ctypes.c_int, p, p]
        lib.oracle_knn_f64.restype = None
        lib.oracle_pair_dist_f64.argtypes = [p, i64, p, i64, ctypes.c_int, p, p, i64, p]
        lib.oracle_pair_dist_f64.restype = None
        _LIB = lib
    return _LIB


def knn_f64(Q: np.ndarray, R: np.ndarray, k: int):
    """Exact top-k by (FP64 squared L2, lower index).  Returns (int64 idx, f64 dist)."""
    Q = np.ascontiguousarray(Q, dtype=np.float32)
    R = np.ascontiguousarray(R, dtype=np.float32)
    nq, D = Q.shape
    nr = R.shape[0]
    idx = np.empty((nq, k), dtype=np.int64)
    dist = np.empty((nq, k), dtype=np.float64)
    _lib().oracle_knn_f64(Q.ctypes.data, nq, D, R.ctypes.data, nr, D, D, k,
                          idx.ctypes.data, dist.ctypes.data)
    return idx, dist


def pair_dist_f64(Q, R, qi, rj):
    Q = np.ascontiguousarray(Q, dtype=np.float32)
    R = np.ascontiguousarray(R, dtype=np.float32)
    qi = np.ascontiguousarray(qi, dtype=np.int64).ravel()
    rj = np.ascontiguousarray(rj, dtype=np.int64).ravel()
    out = np.empty(qi.shape[0], dtype=np.float64)
    _lib().oracle_pair_dist_f64(Q.ctypes.data, Q.shape[1], R.ctypes.data, R.shape[1],
                                Q.shape[1], qi.ctypes.data, rj.ctypes.data, qi.shape[0],
                                out.ctypes.data)
    return out


def knn_numpy_small(Q, R, k):
    """Independent numpy brute force (for pinning the C loop on tiny inputs)."""
    Q = np.asarray(Q, dtype=np.float64)
    R = np.asarray(R, dtype=np.float64)
    d = ((Q[:, None, :] - R[None, :, :]) ** 2).sum(-1)
    idx = np.empty((Q.shape[0], k), dtype=np.int64)
    for i in range(Q.shape[0]):
        order = np.lexsort((np.arange(R.shape[0]), d[i]))[:k]
        idx[i] = order
    return idx, np.take_along_axis(d, idx, 1)


def fix_knn_query(indices: np.ndarray, distances: np.ndarray, ref_idx: np.ndarray):
    """``ann.py:31-52``."""
    fixed_ind, fixed_dist = indices.copy()[:, 1:], distances.copy()[:, 1:]
    mis = indices[:, 0] != ref_idx
    n_mis = int(mis.sum())
    for n in np.where(mis)[0]:
        i, j, kk = ref_idx[n], indices[n], distances[n]
        p = np.where(j == i)[0]
        if len(p) > 0:
            p = p[0]
            j = np.concatenate([j[:p], j[p + 1:]])
            kk = np.concatenate([kk[:p], kk[p + 1:]])
        else:
            j, kk = j[:-1], kk[:-1]
        fixed_ind[n] = j
        fixed_dist[n] = kk
    return fixed_ind, fixed_dist, n_mis


def self_query_knn(E: np.ndarray, k: int, query_rows=None):
    """``knn_utils.py:23-76`` on an exact index: k+1 query, self repair, recall %."""
    rows = np.arange(E.shape[0]) if query_rows is None else np.asarray(query_rows)
    idx, dist = knn_f64(E[rows], E, k + 1)
    fi, fd, nm = fix_knn_query(idx, dist, rows)
    recall = 100.0 * (E.shape[0] - nm) / E.shape[0]
    return fi, fd, nm, recall
