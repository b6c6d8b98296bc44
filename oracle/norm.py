"""FP64 restatement of scarf's RNA normalisation and feature statistics (oracle only).

References (scarf source tree):
* ``assay.py:54-64``      norm_lib_size_log: log1p(sf * counts / libsize)
* ``assay.py:41-51``      norm_lib_size (log_transform=False)
* ``assay.py:770-778``    RNAassay size factor sf (default 1000, attr ``size_factor``)
* ``assay.py:812-828``    renormalize_subset: libsize = row sum over the selected
                          features, zero library sizes become 1
* ``graph_datastore.py:767-796``  mu = clean_array(mean), sigma = clean_array(std, 1)
* ``utils.py:143-154``    clean_array: NaN -> 0, +-inf -> 0, then 0 -> fill_val
* ``ann.py:191-192``      transform_z: (x - mu) / sigma  (no clipping)
"""
from __future__ import annotations

from typing import Optional

import numpy as np


def libsize(counts_dense: np.ndarray, feat_idx: Optional[np.ndarray] = None,
            renormalize_subset: bool = True,
            n_counts: Optional[np.ndarray] = None) -> np.ndarray:
    """Per-cell library size.  ``assay.py:818-825``."""
    if renormalize_subset:
        sub = counts_dense if feat_idx is None else counts_dense[:, feat_idx]
        a = sub.sum(axis=1).astype(np.float64)
        a[a == 0] = 1
        return a
    if n_counts is None:
        raise ValueError("renormalize_subset=False needs the per-cell total counts")
    return np.asarray(n_counts, dtype=np.float64)


def normed(counts_dense: np.ndarray, sf: float = 1000.0, log_transform: bool = True,
           feat_idx: Optional[np.ndarray] = None, renormalize_subset: bool = True,
           n_counts: Optional[np.ndarray] = None) -> np.ndarray:
    """Dense FP64 normalised matrix over the selected features.  ``assay.py:54-64``."""
    ls = libsize(counts_dense, feat_idx, renormalize_subset, n_counts)
    sub = counts_dense if feat_idx is None else counts_dense[:, feat_idx]
    v = (float(sf) * sub.astype(np.float64)) / ls.reshape(-1, 1)
    return np.log1p(v) if log_transform else v


def clean_array(x: np.ndarray, fill_val: float = 0) -> np.ndarray:
    """``utils.py:143-154``."""
    x = np.nan_to_num(np.asarray(x, dtype=np.float64), copy=True)
    x[(x == np.inf) | (x == -np.inf)] = 0
    x[x == 0] = fill_val
    return x


def mu_sigma(x: np.ndarray):
    """Two-pass FP64 mean and population std (ddof=0).  ``graph_datastore.py:767-796``."""
    mu = clean_array(x.mean(axis=0))
    sigma = clean_array(x.std(axis=0), 1)
    return mu, sigma


def transform_z(x: np.ndarray, mu: np.ndarray, sigma: np.ndarray) -> np.ndarray:
    """``ann.py:191-192``; the optional clip is this build's switch (scarf has none)."""
    return (x - mu) / sigma


def transform_z_clipped(x, mu, sigma, clip: Optional[float]):
    z = transform_z(x, mu, sigma)
    if clip is not None:
        z = np.clip(z, -clip, clip)
    return z


def csr_libsize(vals, cols, rowptr, n_cols, feat_idx=None):
    """``libsize`` over the selected features of CSR counts (FP64 integer sums, exact)."""
    n = rowptr.shape[0] - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    keep = _feature_mask(cols, n_cols, feat_idx)
    ls = np.bincount(rows[keep], weights=np.asarray(vals, dtype=np.float64)[keep], minlength=n)
    ls[ls == 0] = 1
    return ls


def _feature_mask(cols, n_cols, feat_idx):
    if feat_idx is None:
        return np.ones(cols.shape[0], dtype=bool)
    sel = np.zeros(n_cols, dtype=bool)
    sel[np.asarray(feat_idx)] = True
    return sel[cols]


def csr_mu_sigma(vals, cols, rowptr, n_cols, sf=1000.0, log_transform=True):
    """``mu_sigma(normed(...))`` for CSR counts over all features, without a dense copy:
    zeros contribute x = 0, so mu_j = sum_nz x / n and the two-pass variance is
    (sum_nz (x - mu_j)^2 + (n - nnz_j) mu_j^2) / n (ddof 0), then clean_array."""
    n = rowptr.shape[0] - 1
    ls = csr_libsize(vals, cols, rowptr, n_cols)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    v = (float(sf) * np.asarray(vals, dtype=np.float64)) / ls[rows]
    x = np.log1p(v) if log_transform else v
    cols = np.asarray(cols, dtype=np.int64)
    s1 = np.bincount(cols, weights=x, minlength=n_cols)
    mu = s1 / n
    nnz_j = np.bincount(cols, minlength=n_cols)
    dev = x - mu[cols]
    m2 = np.bincount(cols, weights=dev * dev, minlength=n_cols) + (n - nnz_j) * mu * mu
    return clean_array(mu), clean_array(np.sqrt(m2 / n), 1)


def csr_z_rows(vals, cols, rowptr, n_cols, rows, mu, sigma, sf=1000.0, clip=None,
               log_transform=True, libsize=None):
    """Dense FP64 z = clip((x - mu) / sigma) of the listed rows of CSR counts (all
    features); ``libsize`` (per row of the whole matrix) defaults to the row sums."""
    rows = np.asarray(rows, dtype=np.int64)
    out = np.zeros((rows.shape[0], n_cols), dtype=np.float64)
    for t, r in enumerate(rows):
        a, b = int(rowptr[r]), int(rowptr[r + 1])
        c = np.asarray(vals[a:b], dtype=np.float64)
        ls = float(libsize[r]) if libsize is not None else (c.sum() or 1.0)
        v = float(sf) * c / ls
        out[t, np.asarray(cols[a:b], dtype=np.int64)] = np.log1p(v) if log_transform else v
    return transform_z_clipped(out, mu, sigma, clip)
