"""Restatement of scarf's kNN edge-weight smoothing (oracle only).

scarf's ``smoothen_dists`` (``knn_utils.py:89-159``) walks the kNN table in
``chunk_size`` row chunks and, per chunk, calls umap-learn's ``smooth_knn_dist`` and
``compute_membership_strengths`` (``knn_utils.py:103``, ``128-134``).  umap-learn is
unpinned in the reference (``requirements.txt:23``) and absent here; its published
algorithm (umap/umap_.py, BSD-3) is restated below with its float32 working precision
(numba locals psum/lo/mid/hi are float32):

* target = log2(k) * bandwidth
* rho_i  = the local_connectivity-th non-zero distance of the row (interpolated)
* sigma_i by at most 64 bisection steps on sum_{j>=1} f(d_ij - rho_i) = target, where
  f(x) = exp(-x / sigma) for x > 0 else 1, stopping when |psum - target| < 1e-5.
  The sum starts at column 1, so with self already removed the nearest neighbour
  is skipped (kept as in the reference).
* sigma_i floored at 1e-3 * mean(row) if rho_i > 0 else 1e-3 * mean(chunk)
* weight = 0 if the neighbour index equals the row's position *within the chunk*,
  1 if d - rho <= 0 or sigma == 0, else exp(-(d - rho) / sigma)

After the chunk loop, zero weights are replaced by ``global_min``: it starts at 1 and
is lowered only by chunks that contain a zero weight, to that chunk's smallest
non-zero weight (``knn_utils.py:144-158``).

Pinned by the reference's fixture ``scarf/tests/datasets/knn_distances.npy`` ->
``knn_weights.npy`` (tolerance 1e-5, ``scarf/tests/test_datastore.py:76-79``).
"""
from __future__ import annotations

import numpy as np

SMOOTH_K_TOLERANCE = 1e-5
MIN_K_DIST_SCALE = 1e-3


def smooth_knn_dist(distances, k, n_iter=64, local_connectivity=1.0, bandwidth=1.0):
    d32 = np.ascontiguousarray(distances, dtype=np.float32)
    n = d32.shape[0]
    target = np.log2(k) * bandwidth
    rho = np.zeros(n, dtype=np.float32)
    result = np.zeros(n, dtype=np.float32)
    mean_distances = np.float32(d32.mean(dtype=np.float64))
    f32 = np.float32
    for i in range(n):
        lo, hi, mid = f32(0.0), f32(np.inf), f32(1.0)
        row = d32[i]
        nz = row[row > 0.0]
        if nz.shape[0] >= local_connectivity:
            index = int(np.floor(local_connectivity))
            interpolation = local_connectivity - index
            if index > 0:
                rho[i] = nz[index - 1]
                if interpolation > SMOOTH_K_TOLERANCE:
                    rho[i] += f32(interpolation * (nz[index] - nz[index - 1]))
            else:
                rho[i] = f32(interpolation * nz[0])
        elif nz.shape[0] > 0:
            rho[i] = np.max(nz)
        for _ in range(n_iter):
            psum = f32(0.0)
            for j in range(1, d32.shape[1]):
                dd = f32(row[j] - rho[i])
                if dd > 0:
                    psum = f32(psum + np.exp(f32(-(dd / mid))))
                else:
                    psum = f32(psum + f32(1.0))
            if abs(float(psum) - target) < SMOOTH_K_TOLERANCE:
                break
            if psum > target:
                hi = mid
                mid = f32((lo + hi) / f32(2.0))
            else:
                lo = mid
                if hi == np.inf:
                    mid = f32(mid * f32(2.0))
                else:
                    mid = f32((lo + hi) / f32(2.0))
        result[i] = mid
        if rho[i] > 0.0:
            m = f32(row.mean(dtype=np.float64))
            if result[i] < MIN_K_DIST_SCALE * m:
                result[i] = f32(MIN_K_DIST_SCALE * m)
        else:
            if result[i] < MIN_K_DIST_SCALE * mean_distances:
                result[i] = f32(MIN_K_DIST_SCALE * mean_distances)
    return result, rho


def compute_membership_strengths(knn_indices, knn_dists, sigmas, rhos):
    """Row-major (i, j) weights as float32; ``i`` is the row's position in the chunk."""
    kd = np.ascontiguousarray(knn_dists, dtype=np.float32)
    n, k = knn_indices.shape
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    cols = np.asarray(knn_indices, dtype=np.int64).ravel()
    x = kd - rhos[:, None]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ex = np.exp(-(x / sigmas[:, None])).astype(np.float32)
    vals = np.where((x <= 0) | (sigmas[:, None] == 0), np.float32(1.0), ex)
    vals = np.where(np.asarray(knn_indices).astype(np.int64) == np.arange(n)[:, None],
                    np.float32(0.0), vals).astype(np.float32)
    return rows, cols, vals.ravel()


def smoothen_dists(z_idx, z_dist, lc=1.0, bw=1.5, chunk_size=1000):
    """``knn_utils.py:89-159``: returns (edges [n*k, 2] uint64, weights [n*k] float64)."""
    n, k = z_idx.shape
    edges = np.zeros((n * k, 2), dtype=np.uint64)
    weights = np.zeros(n * k, dtype=np.float64)
    last_row, val_counts = 0, 0
    null_idx = []
    global_min = 1.0
    for s in range(0, n, chunk_size):
        e = min(s + chunk_size, n)
        ki, kv = z_idx[s:e], np.asarray(z_dist[s:e]).astype(np.float32)
        sig, rho = smooth_knn_dist(kv, k, local_connectivity=lc, bandwidth=bw)
        rows, cols, vals = compute_membership_strengths(ki, kv, sig, rho)
        rows = rows + last_row
        a, b = val_counts, val_counts + len(rows)
        last_row = rows[-1] + 1
        val_counts = b
        edges[a:b, 0] = rows
        edges[a:b, 1] = cols
        weights[a:b] = vals
        nidx = vals == 0
        if nidx.sum() > 0:
            min_val = vals[~nidx].min()
            if min_val < global_min:
                global_min = min_val
        null_idx.extend(nidx)
    weights[np.asarray(null_idx, dtype=bool)] = global_min
    return edges, weights
