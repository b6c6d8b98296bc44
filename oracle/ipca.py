"""FP64 restatement of scarf's PCA fit (oracle only).

scarf fits ``sklearn.decomposition.IncrementalPCA(n_components=dims + 1)`` block by
block and keeps ``components_[:-1].T`` as loadings (``ann.py:207-256``).  The blocks
are fed in scarf's order: the first block with at least dims+1 rows is held back
("end_reservoir") and fitted last, blocks with fewer rows are carried over and
stacked onto the next one, and a block whose SVD raises ``LinAlgError`` is carried
over (``ann.py:215-255``).

``IPCAGram`` restates ``IncrementalPCA.partial_fit``
(scikit-learn ``decomposition/_incremental_pca.py:257-385``, pinned by the reference at
scikit-learn 1.6.1 in ``requirements.txt:16``) without forming the augmented matrix:
the SVD of ``[diag(S) V ; X - xbar_c ; m_c]`` is replaced by the eigendecomposition of
its Gram matrix ``V^T S^2 V + (X - xbar_c)^T (X - xbar_c) + m_c m_c^T``.  Right
singular vectors and eigenvectors coincide, singular values are the square roots of
the eigenvalues, and the sign rule is scikit-learn's ``svd_flip(u_based_decision=False)``
(``utils/extmath.py:895-950``): each component's largest-magnitude entry is positive.
The running mean/variance follow ``_incremental_mean_and_var``
(``utils/extmath.py:1064-1190``).
"""
from __future__ import annotations

from typing import Iterable, List, Optional

import numpy as np


class LinAlgFailure(np.linalg.LinAlgError):
    pass


def incremental_mean_and_var(X, last_mean, last_variance, last_sample_count):
    """``sklearn/utils/extmath.py:1064-1190`` (FP64, no NaN handling needed here)."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    last_count = float(last_sample_count)
    last_sum = np.asarray(last_mean, dtype=np.float64) * last_count
    new_sum = X.sum(axis=0)
    updated_count = last_count + n
    updated_mean = (last_sum + new_sum) / updated_count
    T = new_sum / n
    temp = X - T
    correction = temp.sum(axis=0)
    new_unnorm_var = (temp ** 2).sum(axis=0) - correction ** 2 / n
    if last_count == 0:
        updated_unnorm_var = new_unnorm_var
    else:
        last_unnorm_var = np.asarray(last_variance, dtype=np.float64) * last_count
        last_over_new = last_count / n
        updated_unnorm_var = (
            last_unnorm_var
            + new_unnorm_var
            + last_over_new / updated_count * (last_sum / last_over_new - new_sum) ** 2
        )
    return updated_mean, updated_unnorm_var / updated_count, updated_count


def svd_flip_rows(Vt: np.ndarray) -> np.ndarray:
    """``svd_flip(u_based_decision=False)``: largest |entry| of each row made positive."""
    idx = np.argmax(np.abs(Vt), axis=1)
    signs = np.sign(Vt[np.arange(Vt.shape[0]), idx])
    return Vt * signs[:, None]


class IPCAGram:
    """``IncrementalPCA`` restated through FP64 Gram matrices and ``eigh``."""

    def __init__(self, n_components: int):
        self.n_components = int(n_components)
        self.n_samples_seen_ = 0
        self.mean_ = 0.0
        self.var_ = 0.0
        self.components_: Optional[np.ndarray] = None
        self.singular_values_: Optional[np.ndarray] = None

    def partial_fit_gram(self, n, batch_mean, centred_gram, col_mean, col_var, n_total):
        """Merge step given the chunk's statistics and centred Gram (the GPU's split)."""
        if self.components_ is None:
            M = centred_gram.copy()
        else:
            mc = np.sqrt((self.n_samples_seen_ / n_total) * n) * (self.mean_ - batch_mean)
            M = (self.components_.T * self.singular_values_ ** 2) @ self.components_
            M += centred_gram
            M += np.outer(mc, mc)
        w, V = np.linalg.eigh(M)
        order = np.argsort(w)[::-1]
        w = np.clip(w[order], 0.0, None)
        Vt = svd_flip_rows(V[:, order].T)
        S = np.sqrt(w)
        nc = self.n_components
        self.n_samples_seen_ = n_total
        self.components_ = Vt[:nc]
        self.singular_values_ = S[:nc]
        self.mean_ = col_mean
        self.var_ = col_var
        self.explained_variance_ = (S ** 2 / (n_total - 1))[:nc]
        self.explained_variance_ratio_ = (S ** 2 / np.sum(col_var * n_total))[:nc]
        return self

    def partial_fit(self, X: np.ndarray):
        """``_incremental_pca.py:257-385`` with the SVD replaced by eigh of the Gram."""
        X = np.asarray(X, dtype=np.float64)
        n, d = X.shape
        if self.n_components > d:
            raise ValueError("n_components=%d invalid for n_features=%d" % (self.n_components, d))
        if self.components_ is None and self.n_components > n:
            raise ValueError("n_components must be <= the first batch's sample count")
        col_mean, col_var, n_total = incremental_mean_and_var(
            X, self.mean_, self.var_, self.n_samples_seen_)
        if self.n_samples_seen_ == 0:
            bm = col_mean
        else:
            bm = X.mean(axis=0)
        Xc = X - bm
        return self.partial_fit_gram(n, bm, Xc.T @ Xc, col_mean, col_var, n_total)


def handle_batch_size(dims: int, n_cells: int, batch_size: int) -> int:
    """``ann.py:173-185``: dims capped by the cell count and by batch_size - 1."""
    if dims > n_cells:
        dims = n_cells
    if dims >= batch_size:
        dims = batch_size - 1
    return dims


def scarf_block_schedule(block_sizes: List[int], dims: int) -> List[List[int]]:
    """The sequence of partial_fit calls of ``ann.py:215-255`` as lists of block ids.

    Each entry is the list of blocks stacked into one partial_fit call (in stacking
    order).  The held-back first block is the last call.  (The LinAlgError carry-over
    cannot be scheduled in advance; ``fit_pca_scarf`` handles it.)
    """
    calls, carry, reservoir = [], [], None
    for b, sz in enumerate(block_sizes):
        cur = carry + [b]
        rows = sum(block_sizes[x] for x in cur)
        carry = []
        if rows < dims + 1:
            carry = cur
            continue
        if reservoir is None:
            reservoir = cur
            continue
        calls.append(cur)
    last = (reservoir or []) + carry
    calls.append(last)
    return calls


def fit_pca_scarf(blocks: Iterable[np.ndarray], dims: int, mu, sigma,
                  use_for_pca: Optional[np.ndarray] = None,
                  disable_scaling: bool = False, fitter=None):
    """``AnnStream._fit_pca`` (``ann.py:207-256``) over dense FP64 normalised blocks.

    Returns ``(loadings [d, dims], fitted IPCA object)``.  ``fitter`` defaults to
    ``IPCAGram(dims + 1)``; tests pass scikit-learn's IncrementalPCA to pin it.
    """
    pca = fitter if fitter is not None else IPCAGram(dims + 1)
    s = 0
    end_reservoir, carry_over = None, None
    for blk in blocks:
        i = np.asarray(blk, dtype=np.float64)
        if use_for_pca is not None:
            e = s + i.shape[0]
            i = i[use_for_pca[s:e]]
            s = e
        if not disable_scaling:
            i = (i - mu) / sigma
        if carry_over is not None:
            i = np.vstack((carry_over, i))
            carry_over = None
        if len(i) < dims + 1:
            carry_over = i
            continue
        if end_reservoir is None:
            end_reservoir = i
            continue
        try:
            pca.partial_fit(i)
        except np.linalg.LinAlgError:
            carry_over = i
    if carry_over is not None:
        i = np.vstack((end_reservoir, carry_over)) if end_reservoir is not None else carry_over
    else:
        i = end_reservoir
    try:
        pca.partial_fit(i)
    except np.linalg.LinAlgError:
        pass
    return pca.components_[:-1, :].T, pca


def exact_pca_loadings(z: np.ndarray, dims: int) -> np.ndarray:
    """Textbook PCA of the whole z matrix (centred SVD, svd_flip on rows)."""
    zc = z - z.mean(axis=0)
    _, _, Vt = np.linalg.svd(zc, full_matrices=False)
    return svd_flip_rows(Vt)[:dims].T
