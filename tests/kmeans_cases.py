"""Seeded inputs for the MiniBatchKMeans checks (shared by the CPU and GPU tests).

Each case is (rows float64 [n x d], n_clusters, batch_size, block, init_size); the blocks
are fed to partial_fit in row order, ``block`` rows at a time, as scarf feeds its chunks.
The clusters are separated well enough that no label is decided by a near tie.
"""
import numpy as np


def blobs(seed, n, d, n_true, scale=5.0):
    r = np.random.RandomState(seed)
    cen = r.normal(scale=scale, size=(n_true, d))
    return cen[r.randint(0, n_true, n)] + r.normal(size=(n, d))


def outliers(seed, n_out=40, n=3000, d=5, first=100):
    """n_out far outliers in the first block: k-means++ seeds one center on each, they never
    gain members, and later steps reassign more than half a block of them at once (the cap
    at 0.5 * batch with scikit-learn's argsort order)."""
    r = np.random.RandomState(seed)
    out = r.normal(scale=1000.0, size=(n_out, d))
    blob = r.normal(size=(n, d)) + r.normal(scale=6.0, size=(4, d))[r.randint(0, 4, n)]
    return np.vstack([out, blob[:first - n_out], blob[first - n_out:]])


CASES = {
    # name: (rows, n_clusters, batch_size, block, init_size, first_block)
    "blobs_k8": lambda: (blobs(0, 600, 5, 8), 8, 100, 100, None, None),
    "blobs_k30": lambda: (blobs(1, 2000, 12, 30), 30, 256, 256, None, None),
    "small_blocks_k40": lambda: (blobs(2, 900, 7, 40), 40, 50, 50, None, None),
    "init_subsample": lambda: (blobs(3, 1500, 10, 20), 20, 50, 300, None, None),
    "explicit_init_size": lambda: (blobs(5, 1200, 9, 25), 25, 200, 200, 80, None),
    "k_equals_first_block": lambda: (blobs(7, 1000, 6, 12), 100, 100, 100, None, None),
    "wide_k150_d70": lambda: (blobs(8, 3000, 70, 40, scale=3.0), 150, 500, 500, None, None),
    "outliers_cap": lambda: (outliers(21), 45, 50, 50, None, 100),
    "two_clusters_big_init": lambda: (blobs(11, 400, 4, 2), 2, 1000, 150, 5000, None),
}


def blocks_of(X, block, first_block=None):
    s = 0
    if first_block:
        yield X[:first_block]
        s = first_block
    while s < X.shape[0]:
        yield X[s:s + block]
        s += block
