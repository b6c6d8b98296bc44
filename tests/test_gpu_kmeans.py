"""GPU MiniBatchKMeans (scarf's seed partitions, ann.py:328-346) against the oracle.

The host makes scikit-learn's random draws in scikit-learn's order, so k-means++ seeds,
reassignments, counts and labels must agree exactly, and centers bit for bit: the center
update is the same chain of roundings.  Distances differ from the oracle's BLAS products
only in rounding, which decides nothing on these inputs (no near ties).
"""
import numpy as np
import pytest

from conftest import say
from kmeans_cases import CASES, blocks_of
from oracle import kmeans as okm

pytestmark = pytest.mark.gpu


def _fit_both(X, k, bs, block, init_size, first, rows, seed=7, compute_labels=True):
    import torch
    from scarf_amd.kmeans import MiniBatchKMeans
    Xh = X.astype(np.float32).astype(np.float64) if rows == "f32" else X
    Xd = torch.from_numpy(Xh.astype(np.float32) if rows == "f32" else Xh).cuda()
    ref = okm.MiniBatchKMeans(n_clusters=k, random_state=seed, batch_size=bs,
                              init_size=init_size)
    gpu = MiniBatchKMeans(n_clusters=k, random_state=seed, batch_size=bs, n_init=3,
                          init_size=init_size, compute_labels=compute_labels)
    s = 0
    for b in blocks_of(Xh, block, first):
        e = s + b.shape[0]
        ref.partial_fit(b)
        gpu.partial_fit(Xd[s:e])
        s = e
    return Xh, Xd, ref, gpu


@pytest.mark.parametrize("rows", ["f32", "f64"])
@pytest.mark.parametrize("case", list(CASES))
def test_kmeans_gpu_matches_oracle(cuda, case, rows):
    X, k, bs, block, init_size, first = CASES[case]()
    say("kmeans", case, rows, X.shape, "k", k, "batch", bs)
    Xh, Xd, ref, gpu = _fit_both(X, k, bs, block, init_size, first, rows)
    np.testing.assert_array_equal(gpu.init_indices_, ref.init_indices_)
    np.testing.assert_array_equal(gpu._counts, ref._counts)
    np.testing.assert_array_equal(gpu.predict(Xd), ref.predict(Xh))
    C = gpu.cluster_centers_
    assert np.array_equal(C, ref.cluster_centers_), np.abs(C - ref.cluster_centers_).max()
    if case == "outliers_cap":
        assert ref.reassign_log and len(ref.reassign_log[0][1]) == 25


def test_kmeans_labels_and_host_rows(cuda):
    """labels_ (compute_labels) are the last block's labels under the updated centers; host
    numpy rows give the same fit as device rows."""
    from scarf_amd.kmeans import MiniBatchKMeans
    X, k, bs, block, init_size, first = CASES["blobs_k30"]()
    say("kmeans labels_ / host rows")
    Xh, Xd, ref, gpu = _fit_both(X, k, bs, block, init_size, first, "f64")
    last = Xh[(Xh.shape[0] - 1) // block * block:]
    np.testing.assert_array_equal(gpu.labels_, ref.predict(last))
    host = MiniBatchKMeans(n_clusters=k, random_state=7, batch_size=bs, n_init=3)
    for b in blocks_of(Xh, block, first):
        host.partial_fit(b)
    assert np.array_equal(host.cluster_centers_, gpu.cluster_centers_)
    with pytest.raises(ValueError):
        host.partial_fit(Xh[:, :3])
    with pytest.raises(ValueError):
        MiniBatchKMeans(n_clusters=50, random_state=0).partial_fit(Xh[:20])
