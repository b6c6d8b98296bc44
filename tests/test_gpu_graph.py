"""GPU graph loading and the on-disk store against scipy's sparse arithmetic (the oracle).

The CSR arrays must be identical: sorted columns, the same entries (explicit zeros kept in
g, zeros produced by the symmetrisation dropped) and bit-identical values, since every
value is one rounding of a + b, a * b or s - p.
"""
import numpy as np
import pytest

from conftest import say
from oracle import graph as og

pytestmark = pytest.mark.gpu


def knn_like(n, k, seed, zero_frac=0.05, mutual=0.4):
    """kNN-shaped edges: row i lists k distinct cells (itself first), weights in (0, 1] with
    some exact zeros, and a fraction of edges mirrored so that both directions exist."""
    rng = np.random.default_rng(seed)
    cols = np.empty((n, k), dtype=np.int64)
    for i in range(n):
        others = rng.choice(n - 1, k - 1, replace=False)
        others[others >= i] += 1
        cols[i] = np.concatenate([[i], others])
    for i in rng.choice(n, int(mutual * n), replace=False):
        j = cols[i, 1]
        if i not in cols[j]:
            cols[j, k - 1] = i
    w = rng.random((n, k))
    w[rng.random((n, k)) < zero_frac] = 0.0
    edges = np.stack([np.repeat(np.arange(n), k), cols.ravel()], axis=1).astype(np.uint64)
    return edges, w.ravel()


def same(got, want):
    assert got.format == want.format
    if got.format == "coo":
        np.testing.assert_array_equal(got.row, want.row)
        np.testing.assert_array_equal(got.col, want.col)
    else:
        np.testing.assert_array_equal(got.indptr, want.indptr)
        np.testing.assert_array_equal(got.indices, want.indices)
    np.testing.assert_array_equal(got.data, want.data)
    assert got.shape == want.shape


@pytest.mark.parametrize("use_k", [None, 5, 1, 40])
@pytest.mark.parametrize("mode", ["plain", "symmetric", "upper"])
def test_graph_csr_matches_scipy(cuda, use_k, mode):
    from scarf_amd.sparse import graph_csr
    n, k = 3000, 11
    edges, w = knn_like(n, k, seed=5)
    sym, up = mode != "plain", mode == "upper"
    say("graph_csr", mode, "use_k", use_k)
    got = graph_csr(edges, w, n, k, use_k=use_k, symmetric=sym, upper_only=up)
    want = og.load_graph(edges, w, n, k, symmetric=sym, upper_only=up, use_k=use_k)
    same(got, want)
    if sym and not up:
        assert (abs(got - got.T) > 0).nnz == 0


def test_graph_csr_single_cell(cuda):
    """One cell whose only neighbour is itself: a 1 x 1 matrix, symmetrised to a + a - a*a."""
    from scarf_amd.sparse import graph_csr
    edges = np.array([[0, 0]], dtype=np.uint64)
    for w in (np.array([0.25]), np.array([0.0])):
        for sym in (False, True):
            same(graph_csr(edges, w, 1, 1, symmetric=sym),
                 og.load_graph(edges, w, 1, 1, symmetric=sym))


def test_graph_csr_duplicates_and_bad_edges(cuda):
    """Two copies of an edge are summed (scipy's coo -> csr); an edge naming a cell outside
    the graph is an error, not a write."""
    from scarf_amd.sparse import graph_csr
    edges, w = knn_like(500, 6, seed=9)
    edges[7] = edges[8]  # row 1 holds (1, c) twice
    say("graph_csr duplicates")
    for sym in (False, True):
        same(graph_csr(edges, w, 500, 6, symmetric=sym), og.load_graph(edges, w, 500, 6,
                                                                        symmetric=sym))
    bad = edges.copy()
    bad[3, 1] = 500
    with pytest.raises(ValueError):
        graph_csr(bad, w, 500, 6)


def test_persistent_store_make_graph_load_graph(cuda, tmp_path):
    """make_graph into a DirStore, then a second session on the same directory: nothing is
    recomputed (the loaded index is reused), the outputs are the same, load_graph reads
    the stored edges, and run_mapping through the reloaded index maps as before."""
    from scarf_amd.datastore import load_graph, make_graph, run_mapping
    from scarf_amd.store import DirStore
    from scarf_amd.synth import generate_csr, make_gene_model

    model = make_gene_model(200, 3, 0.12, 17)
    csr = generate_csr(1500, model, 17)
    say("make_graph into a DirStore")
    s1 = DirStore(str(tmp_path / "z"))
    a1 = make_graph(csr, None, dims=10, k=7, batch_size=500, store=s1, return_ann_object=True)
    knn_loc, g_loc = s1["latest_knn"], s1["latest_graph"]
    s2 = DirStore(str(tmp_path / "z"))
    say("second session")
    a2 = make_graph(csr, None, dims=10, k=7, batch_size=500, store=s2, return_ann_object=True)
    ann_key = knn_loc.rsplit("/", 1)[0] + "/ann_idx"
    assert s2._loaded.get(ann_key) is a2.annIdx  # the saved index, not a rebuilt one
    for key in (knn_loc + "/indices", knn_loc + "/distances", g_loc + "/edges",
                g_loc + "/weights"):
        np.testing.assert_array_equal(s1[key], s2[key])
    kl = s2["RNA/normed__I__hvgs/reduction__pca__10__I/latest_kmeans"]
    np.testing.assert_array_equal(s2[kl + "/cluster_labels"], a1.clusterLabels)
    want = og.load_graph(s1[g_loc + "/edges"], s1[g_loc + "/weights"], 1500, 7,
                         symmetric=True)
    same(load_graph(s2, symmetric=True), want)
    ids = [f"g{i}" for i in range(200)]
    run_mapping(a1, csr, ids, ids, "q", save_k=3, store=s1)
    run_mapping(a2, csr, ids, ids, "q2", save_k=3, store=s2)
    np.testing.assert_array_equal(s1["RNA/projections/q/indices"],
                                  s2["RNA/projections/q2/indices"])
