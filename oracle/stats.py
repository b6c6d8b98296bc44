"""ctypes wrappers of ``stats_ref.c``: FP64 normalisation statistics and the z Gram of
CSR counts at full size (oracle only; the same arithmetic as ``oracle.norm``, pinned
against it in tests/test_oracle.py).

References (scarf source tree): ``assay.py:54-64, 818-823`` (library size, log1p),
``graph_datastore.py:767-796`` and ``utils.py:143-154`` (mu, sigma, clean_array),
``ann.py:191-192`` (z), ``ann.py:207-256`` with scikit-learn
``_incremental_pca.py:257-385`` (the Gram IncrementalPCA's first partial_fit
decomposes).
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
        path = os.path.join(_HERE, "_build", "liboracle_stats.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        lib = ctypes.CDLL(path)
        p, i64, f64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
        lib.oracle_csr_stats.argtypes = [p, p, p, i64, i64, f64, i32, p, p, p]
        lib.oracle_csr_stats.restype = None
        lib.oracle_csr_zgram.argtypes = [p, p, p, i64, i64, p, f64, i32, p, p, f64, p, p]
        lib.oracle_csr_zgram.restype = None
        _LIB = lib
    return _LIB


def _csr(vals, cols, rowptr):
    return (np.ascontiguousarray(vals).view(np.uint32),
            np.ascontiguousarray(cols, dtype=np.int32),
            np.ascontiguousarray(rowptr, dtype=np.int64))


def csr_stats(vals, cols, rowptr, n_cols, sf=1000.0, log_transform=True):
    """(libsize, mu, sigma) of CSR counts over all features, FP64."""
    v, c, rp = _csr(vals, cols, rowptr)
    n = rp.shape[0] - 1
    ls = np.empty(n, np.float64)
    mu = np.empty(n_cols, np.float64)
    sg = np.empty(n_cols, np.float64)
    _lib().oracle_csr_stats(v.ctypes.data, c.ctypes.data, rp.ctypes.data, n, n_cols,
                            float(sf), int(log_transform), ls.ctypes.data, mu.ctypes.data,
                            sg.ctypes.data)
    return ls, mu, sg


def csr_zgram(vals, cols, rowptr, n_cols, libsize, mu, sigma, sf=1000.0, clip=None,
              log_transform=True):
    """(Z^T Z [d, d], column sums of z [d]) with z = clip((x - mu) / sigma), FP64."""
    v, c, rp = _csr(vals, cols, rowptr)
    n = rp.shape[0] - 1
    ls = np.ascontiguousarray(libsize, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    sg = np.ascontiguousarray(sigma, dtype=np.float64)
    G = np.empty((n_cols, n_cols), np.float64)
    cs = np.empty(n_cols, np.float64)
    _lib().oracle_csr_zgram(v.ctypes.data, c.ctypes.data, rp.ctypes.data, n, n_cols,
                            ls.ctypes.data, float(sf), int(log_transform), mu.ctypes.data,
                            sg.ctypes.data, float(clip) if clip else 0.0, G.ctypes.data,
                            cs.ctypes.data)
    return G, cs


def exact_pca(G, colsum, n, n_components):
    """Top components of the centred Gram M = G - colsum colsum^T / n (the first
    partial_fit's matrix): (eigenvalues descending, components [n_components, d]) with
    scikit-learn's sign rule (largest |entry| of each component positive,
    ``utils/extmath.py:895-950``)."""
    import scipy.linalg as sla
    d = G.shape[0]
    M = G - np.outer(colsum, colsum) / n
    M = 0.5 * (M + M.T)
    w, V = sla.eigh(M, subset_by_index=[d - n_components, d - 1], driver="evr")
    w, Vt = w[::-1].copy(), V[:, ::-1].T.copy()
    idx = np.argmax(np.abs(Vt), axis=1)
    Vt *= np.sign(Vt[np.arange(Vt.shape[0]), idx])[:, None]
    return w, Vt
