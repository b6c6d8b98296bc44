"""Restatement of scarf's MiniBatchKMeans seed partitions (oracle only).

scarf's ``AnnStream._fit_kmeans`` (``ann.py:328-346``) builds
``sklearn.cluster.MiniBatchKMeans(n_clusters=nClusters, random_state=randState,
batch_size=batchSize, n_init=3)``, calls ``partial_fit(reducer(block))`` for every block
and then ``predict(reducer(block))`` for every block; ``graph_datastore.py:958-975`` saves
``cluster_centers_`` and the labels.  The reference pins scikit-learn 1.6.1
(``requirements.txt:16``); this restates the installed scikit-learn 1.7.2, dense input,
unit sample weights, ``init="k-means++"``:

* ``partial_fit``        ``sklearn/cluster/_kmeans.py:2203-2303``
* ``_check_params_vs_input`` (batch and init sizes) ``_kmeans.py:1922-1944``
* ``_init_centroids``    ``_kmeans.py:1003-1020`` (``randint`` subsample when init_size < n)
* ``_kmeans_plusplus``   ``_kmeans.py:174-272``
* ``_labels_inertia``    ``_kmeans.py:755-822`` -> ``_k_means_lloyd.pyx:172-215``
  (argmin of ||c||^2 - 2 x.c, first index on ties)
* ``_mini_batch_step``   ``_kmeans.py:1557-1675`` -> ``_k_means_minibatch.pyx:59-108``
  (per center: c*w, + x per member in row order, * (1 / (w + count)))
* ``_random_reassign``   ``_kmeans.py:2029-2043``
* ``predict``            ``_kmeans.py:1066-1098``

Pinned in ``tests/test_oracle.py`` against scikit-learn's own ``MiniBatchKMeans`` and
``kmeans_plusplus`` (seeds, labels, counts and centers identical bit for bit) on inputs
without near ties, including the init_size subsample and the reassignment cap, and on a
hand-computed update chain.
Sums whose order scikit-learn leaves to BLAS are written here with numpy's operators; the
two agree to rounding, which changes no decision on such inputs.
"""
from __future__ import annotations

from typing import Optional

import numpy as np


def row_norms_sq(X: np.ndarray) -> np.ndarray:
    return np.einsum("ij,ij->i", X, X)


def sq_dists(A: np.ndarray, B: np.ndarray, b_norms: np.ndarray) -> np.ndarray:
    """``_euclidean_distances(A, B, Y_norm_squared=b_norms, squared=True)`` for float64
    (``sklearn/metrics/pairwise.py``): -2 A.B^T + ||a||^2 + ||b||^2, clipped at 0."""
    d = -2 * (A @ B.T)
    d += row_norms_sq(A)[:, None]
    d += b_norms[None, :]
    np.maximum(d, 0, out=d)
    return d


def kmeans_plusplus(X: np.ndarray, n_clusters: int, x_norms: np.ndarray,
                    rs: np.random.RandomState, n_local_trials: Optional[int] = None):
    """Greedy k-means++ (``_kmeans.py:174-272``), unit weights.  Returns (centers, indices)."""
    n, d = X.shape
    if n_local_trials is None:
        n_local_trials = 2 + int(np.log(n_clusters))
    w = np.ones(n)
    centers = np.empty((n_clusters, d))
    indices = np.full(n_clusters, -1, dtype=np.int64)
    cid = rs.choice(n, p=w / w.sum())
    centers[0] = X[cid]
    indices[0] = cid
    closest = sq_dists(X[cid][None, :], X, x_norms)
    pot = closest @ w
    for c in range(1, n_clusters):
        r = rs.uniform(size=n_local_trials) * pot
        cand = np.searchsorted(np.cumsum(w * closest, dtype=np.float64), r)
        np.clip(cand, None, closest.size - 1, out=cand)
        dc = sq_dists(X[cand], X, x_norms)
        np.minimum(closest, dc, out=dc)
        cpot = dc @ w.reshape(-1, 1)
        best = np.argmin(cpot)
        pot = cpot[best]
        closest = dc[best]
        centers[c] = X[cand[best]]
        indices[c] = cand[best]
    return centers, indices


def assign(X: np.ndarray, centers: np.ndarray) -> np.ndarray:
    """Nearest center, first index on ties (``_k_means_lloyd.pyx:172-215``)."""
    score = row_norms_sq(centers)[None, :] - 2.0 * (X @ centers.T)
    return np.argmin(score, axis=1).astype(np.int32)


def minibatch_update(X: np.ndarray, centers: np.ndarray, weight_sums: np.ndarray,
                     labels: np.ndarray) -> None:
    """In-place ``_minibatch_update_dense`` with unit weights
    (``_k_means_minibatch.pyx:59-108``): one sequential chain per (center, feature)."""
    for c in range(centers.shape[0]):
        members = np.flatnonzero(labels == c)
        if members.size == 0:
            continue
        acc = centers[c] * weight_sums[c]
        for m in members:
            acc = acc + X[m]
        weight_sums[c] += float(members.size)
        alpha = 1.0 / weight_sums[c]
        centers[c] = acc * alpha


class MiniBatchKMeans:
    """``partial_fit`` / ``predict`` of scikit-learn's MiniBatchKMeans (dense, unit weights)."""

    def __init__(self, n_clusters: int = 8, *, batch_size: int = 1024, random_state=None,
                 init_size: Optional[int] = None, reassignment_ratio: float = 0.01,
                 n_init=3):
        self.n_clusters = int(n_clusters)
        self.batch_size = int(batch_size)
        self.random_state = random_state
        self.init_size = init_size
        self.reassignment_ratio = float(reassignment_ratio)
        self.n_steps_ = 0
        self.reassign_log = []  # (step, cluster indices, chosen rows) per reassignment

    def _rs(self):
        if not hasattr(self, "_random_state"):
            rs = self.random_state
            if rs is None or isinstance(rs, (int, np.integer)):
                rs = np.random.RandomState(rs)
            self._random_state = rs
        return self._random_state

    def partial_fit(self, X: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float64)
        rs = self._rs()
        xn = row_norms_sq(X)
        if not hasattr(self, "cluster_centers_"):
            n = X.shape[0]
            if n < self.n_clusters:
                raise ValueError(f"n_samples={n} should be >= n_clusters={self.n_clusters}.")
            self._batch_size = min(self.batch_size, n)
            init_size = self.init_size
            if init_size is None:
                init_size = 3 * self._batch_size
                if init_size < self.n_clusters:
                    init_size = 3 * self.n_clusters
            elif init_size < self.n_clusters:
                init_size = 3 * self.n_clusters
            init_size = min(init_size, n)
            Xi, xni, ii = X, xn, None
            if init_size < n:
                ii = rs.randint(0, n, init_size)
                Xi, xni = X[ii], xn[ii]
            self.cluster_centers_, idx = kmeans_plusplus(Xi, self.n_clusters, xni, rs)
            self.init_indices_ = idx if ii is None else ii[idx]
            self._counts = np.zeros(self.n_clusters)
            self._n_since_last_reassign = 0
        # _random_reassign() is evaluated before the step runs
        self._n_since_last_reassign += self._batch_size
        reassign = False
        if (self._counts == 0).any() or self._n_since_last_reassign >= 10 * self.n_clusters:
            self._n_since_last_reassign = 0
            reassign = True
        labels = assign(X, self.cluster_centers_)
        minibatch_update(X, self.cluster_centers_, self._counts, labels)
        if reassign and self.reassignment_ratio > 0:
            ws = self._counts
            to_reassign = ws < self.reassignment_ratio * ws.max()
            if to_reassign.sum() > 0.5 * X.shape[0]:
                keep = np.argsort(ws)[int(0.5 * X.shape[0]):]
                to_reassign[keep] = False
            n_re = to_reassign.sum()
            if n_re:
                new = rs.choice(X.shape[0], replace=False, size=n_re)
                self.cluster_centers_[to_reassign] = X[new]
                self.reassign_log.append((self.n_steps_, np.flatnonzero(to_reassign), new))
            ws[to_reassign] = np.min(ws[~to_reassign])
        self.n_steps_ += 1
        return self

    def predict(self, X: np.ndarray) -> np.ndarray:
        return assign(np.asarray(X, dtype=np.float64), self.cluster_centers_)
