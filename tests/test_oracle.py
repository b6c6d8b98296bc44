"""Pins the CPU oracle against the reference's fixtures, scikit-learn and closed forms.

Each check is chosen so that a plausible slip in the oracle (wrong library size,
ddof, log vs log1p, sign rule, block order, tie order, chunk-local self index,
the global-minimum rule, the score normalisation) fails it.
"""
import math
import os

import numpy as np
import pytest

from oracle import ipca, knn, mapping, norm, smooth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- normalisation (assay.py:54-64, 812-828; graph_datastore.py:767-796) ----------
TOY = np.array([[5, 0, 0, 2], [3, 3, 0, 7], [3, 3, 0, 7], [0, 0, 0, 0]], dtype=np.uint32)


def test_norm_subset_libsize_literal():
    x = norm.normed(TOY, sf=1000, feat_idx=np.array([0, 3]))
    # row 0: library size over features {0, 3} = 7
    assert x[0, 0] == pytest.approx(6.572682063274631, rel=1e-14)  # ln(1 + 5000/7)
    assert x[0, 1] == pytest.approx(5.6584861997410245, rel=1e-14)  # ln(1 + 2000/7)
    # row 1: subset library size 10 (not the full 13)
    assert x[1, 1] == pytest.approx(6.55250788703459, rel=1e-14)  # ln(1 + 7000/10)
    # all-zero row: library size 0 -> 1, values 0
    assert np.all(x[3] == 0)


def test_norm_no_log_and_full_counts():
    x = norm.normed(TOY, sf=1000, log_transform=False, renormalize_subset=False,
                    n_counts=np.array([7.0, 13.0, 13.0, 1.0]))
    assert x[1, 3] == pytest.approx(7000 / 13, rel=1e-15)


def test_mu_sigma_ddof0_and_clean():
    x = np.array([[1.0, 5.0, 0.0], [3.0, 5.0, 0.0]])
    mu, sigma = norm.mu_sigma(x)
    np.testing.assert_array_equal(mu, [2.0, 5.0, 0.0])
    np.testing.assert_array_equal(sigma, [1.0, 1.0, 1.0])  # ddof=0; zero std -> 1
    c = norm.clean_array(np.array([np.nan, np.inf, -np.inf, 0.0, 2.0]), 7)
    # nan_to_num runs first, so +-inf become +-DBL_MAX and survive (utils.py:150-151)
    big = np.finfo(np.float64).max
    np.testing.assert_array_equal(c, [7, big, -big, 7, 2])


def test_z_moments():
    rng = np.random.default_rng(0)
    x = rng.gamma(1.0, 1.0, (500, 8))
    mu, sigma = norm.mu_sigma(x)
    z = norm.transform_z(x, mu, sigma)
    np.testing.assert_allclose(z.mean(0), 0, atol=1e-12)
    np.testing.assert_allclose(z.var(0), 1, rtol=1e-12)


# ---- incremental PCA (ann.py:207-256 + scikit-learn IncrementalPCA) --------------
def _sk_ipca(blocks, nc):
    from sklearn.decomposition import IncrementalPCA
    p = IncrementalPCA(n_components=nc)
    for b in blocks:
        p.partial_fit(b.copy())
    return p


def test_ipca_gram_matches_sklearn():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((700, 30)) @ rng.standard_normal((30, 30)) + rng.standard_normal(30)
    blocks = [X[i:i + 160] for i in range(0, 700, 160)]
    sk = _sk_ipca(blocks, 9)
    me = ipca.IPCAGram(9)
    for b in blocks:
        me.partial_fit(b)
    np.testing.assert_allclose(me.components_, sk.components_, atol=1e-10)
    np.testing.assert_allclose(me.singular_values_, sk.singular_values_, rtol=1e-11)
    np.testing.assert_allclose(me.mean_, sk.mean_, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(me.var_, sk.var_, rtol=1e-11)
    np.testing.assert_allclose(me.explained_variance_ratio_, sk.explained_variance_ratio_,
                               rtol=1e-10)


def test_ipca_single_chunk_is_exact_pca():
    rng = np.random.default_rng(2)
    z = rng.standard_normal((300, 12)) @ rng.standard_normal((12, 12))
    L, _ = ipca.fit_pca_scarf([z], 5, 0.0, 1.0)
    np.testing.assert_allclose(L, ipca.exact_pca_loadings(z, 5), atol=1e-10)


def test_ipca_sign_rule():
    Vt = np.array([[0.1, -0.9, 0.3], [-0.2, 0.1, 0.05]])
    f = ipca.svd_flip_rows(Vt)
    assert f[0, 1] > 0 and f[1, 0] > 0
    np.testing.assert_array_equal(np.abs(f), np.abs(Vt))


def test_scarf_block_order():
    # block 0 (3 rows) < dims+1=5 is carried onto block 1 -> held back; block 2 fitted;
    # block 3 (2 rows) carried, stacked after the held-back block in the last call
    calls = ipca.scarf_block_schedule([3, 10, 10, 2], 4)
    assert calls == [[2], [0, 1, 3]]


@pytest.mark.parametrize("name", ["ipca_small", "ipca_blocks", "ipca_wide"])
def test_ipca_golden(name):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    from scarf_amd.synth import CSR
    csr = CSR(g["vals"], g["cols"], g["rowptr"], int(g["n_cols"]))
    counts = csr.to_dense()
    x = norm.normed(counts, sf=float(g["sf"]))
    mu, sigma = norm.mu_sigma(x)
    np.testing.assert_allclose(mu, g["mu"], rtol=1e-13)
    np.testing.assert_allclose(sigma, g["sigma"], rtol=1e-13)
    ch = int(g["chunk"])
    blocks = [x[s:s + ch] for s in range(0, x.shape[0], ch)]
    L, st = ipca.fit_pca_scarf(blocks, int(g["dims"]), mu, sigma)
    np.testing.assert_allclose(L, g["loadings"], atol=1e-9)
    np.testing.assert_allclose(st.explained_variance_ratio_, g["explained_variance_ratio"],
                               rtol=1e-9)


# ---- kNN (hnswlib 'l2' semantics, ann.py:31-52) ----------------------------------
def test_knn_c_matches_numpy():
    rng = np.random.default_rng(3)
    Q = rng.standard_normal((40, 6)).astype(np.float32)
    R = rng.standard_normal((90, 6)).astype(np.float32)
    i1, d1 = knn.knn_f64(Q, R, 7)
    i2, d2 = knn.knn_numpy_small(Q, R, 7)
    np.testing.assert_array_equal(i1, i2)
    np.testing.assert_allclose(d1, d2, rtol=1e-13)


def test_knn_ties_lower_index():
    R = np.array([[1, 0], [0, 0], [1, 0], [0, 0], [5, 5]], dtype=np.float32)
    Q = np.array([[0, 0]], dtype=np.float32)
    i, d = knn.knn_f64(Q, R, 4)
    np.testing.assert_array_equal(i[0], [1, 3, 0, 2])
    np.testing.assert_array_equal(d[0], [0, 0, 1, 1])


def test_fix_knn_query_cases():
    ind = np.array([[5, 1, 2, 3], [7, 8, 6, 9], [1, 2, 3, 4]])
    dist = np.arange(12, dtype=float).reshape(3, 4)
    fi, fd, nm = knn.fix_knn_query(ind, dist, np.array([5, 6, 0]))
    np.testing.assert_array_equal(fi, [[1, 2, 3], [7, 8, 9], [1, 2, 3]])
    np.testing.assert_array_equal(fd, [[1, 2, 3], [4, 5, 7], [8, 9, 10]])
    assert nm == 2


def test_self_query_duplicates():
    E = np.zeros((6, 3), dtype=np.float32)
    E[4:] = 1.0
    fi, fd, nm, rec = knn.self_query_knn(E, 2)
    # rows 0..3 identical: their k+1=3 nearest are 0,1,2 by index, so row 3 does not
    # find itself (last column dropped); rows 1, 2, 3 and 5 do not have themselves in
    # column 0 and count as misses (ann.py:34-35)
    np.testing.assert_array_equal(fi[0], [1, 2])
    np.testing.assert_array_equal(fi[2], [0, 1])
    np.testing.assert_array_equal(fi[3], [0, 1])
    np.testing.assert_array_equal(fi[5], [4, 0])
    assert nm == 4 and rec == pytest.approx(100 * 2 / 6)


# ---- smooth_knn_dist (knn_utils.py:89-159) ---------------------------------------
def test_smooth_reference_fixture():
    ki = np.load(os.path.join(GOLD, "knn_indices.npy"))
    kd = np.load(os.path.join(GOLD, "knn_distances.npy"))
    kw = np.load(os.path.join(GOLD, "knn_weights.npy"))
    edges, w = smooth.smoothen_dists(ki, kd, lc=1.0, bw=1.5, chunk_size=1000)
    assert np.abs(w - kw).max() < 1e-5  # reference test tolerance, two-sided here
    np.testing.assert_array_equal(edges[:, 1], ki.ravel())
    np.testing.assert_array_equal(edges[:, 0], np.repeat(np.arange(808), 11))


def test_smooth_vectorised_matches_scalar():
    """The row-vectorised restatement timed as the CPU baseline equals the scalar one."""
    from oracle.baseline import smooth_knn_dist_vec
    ki = np.load(os.path.join(GOLD, "knn_indices.npy"))
    kd = np.load(os.path.join(GOLD, "knn_distances.npy"))
    s1, r1 = smooth.smooth_knn_dist(kd, 11, local_connectivity=1.0, bandwidth=1.5)
    s2, r2 = smooth_knn_dist_vec(kd, 11, local_connectivity=1.0, bandwidth=1.5)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_allclose(s1, s2, rtol=1e-5)


def test_smooth_chunk_local_self_and_global_min():
    # chunk_size 2.  Chunk 1 = rows 2, 3: row 2 lists neighbour 0 and row 3 lists
    # neighbour 1, each equal to the row's position *inside the chunk* -> weight 0.
    # Chunk 0 has no such entry but holds the smallest weight overall (a far
    # neighbour); it must not lower global_min, which only zero-holding chunks set.
    idx = np.array([[1, 2, 3], [0, 2, 3], [0, 3, 1], [2, 1, 0]])
    d = np.array([[1.0, 2.0, 30.0], [1.0, 2.0, 3.0], [0.5, 1.0, 1.5], [1.0, 1.2, 1.3]])
    _, w = smooth.smoothen_dists(idx, d, lc=1.0, bw=1.0, chunk_size=2)
    w = w.reshape(4, 3)
    raw_c1 = [w[2, 1], w[2, 2], w[3, 0], w[3, 2]]
    gmin = min(raw_c1)
    assert w[2, 0] == gmin and w[3, 1] == gmin
    assert w[0].min() < gmin  # chunk 0 minimum is smaller, and unused
    assert np.all(w > 0)


# ---- mapping scores (mapping_datastore.py:211-285) -------------------------------
def test_mapping_score_literal():
    idx = np.array([[0, 1], [1, 2]])
    dist = np.array([[0.0, math.e - 1], [0.0, math.e - 1]])
    (_, ms), = mapping.mapping_scores(idx, dist, 4)
    # w(0) = 1, w(e - 1) = 1/2: sums [1, 1.5, 0.5, 0]; multiplier 1000 / (2 cells * 2) = 250
    np.testing.assert_allclose(ms, np.log1p([250.0, 375.0, 125.0, 0.0]), rtol=1e-15)


def test_mapping_score_reference_fixture():
    """scarf's selfmap (save_k=3) finds each cell itself first, then its two nearest
    graph neighbours; the fixture cell_attributes.csv:mapping_scores was written by a
    version that divided by the number of target cells only, so it equals the current
    formula (which also divides by k) at multiplier 1000 * k.  (The reference's own
    test compares one-sidedly, so it passes either way.)"""
    import csv
    ki = np.load(os.path.join(GOLD, "knn_indices.npy"))
    kd = np.load(os.path.join(GOLD, "knn_distances.npy"))
    fx = np.load(os.path.join(GOLD, "mapping_scores.npy"))
    n = ki.shape[0]
    idx = np.stack([np.arange(n), ki[:, 0], ki[:, 1]], 1)
    dist = np.stack([np.zeros(n), kd[:, 0], kd[:, 1]], 1)
    (_, ms), = mapping.mapping_scores(idx, dist, n, multiplier=1000 * 3)
    assert np.abs(ms - fx).max() < 1e-12


def test_align_features():
    src = ["a", "b", "c", "d"]
    tgt = ["c", "x", "a"]
    t_idx = mapping.order_features(src, tgt)
    np.testing.assert_array_equal(t_idx, [2, -1, 0, -1])
    counts = np.array([[4, 9, 1]], dtype=np.uint32)  # c=4, x=9, a=1
    out = mapping.align_normed(counts, t_idx, sf=10)
    # library size over present features only: 4 + 1 = 5
    np.testing.assert_allclose(out[0], [np.log1p(10 * 1 / 5), 1.0, np.log1p(10 * 4 / 5), 1.0])


def test_target_classes_oracle_hand_cases():
    """Label transfer (mapping_datastore.py:340-366) on rows worked by hand."""
    from oracle.mapping import target_classes
    groups = np.array(["A", "B", "A", "C", "B"])
    idx = np.array([[0, 1, 2, 3],      # classes A B A | last column ignored
                    [1, 4, 1, 0],      # B B B
                    [0, 1, 3, 2],      # A B C
                    [0, 0, 0, 0]])
    d = np.array([[1.0, 2.0, 3.0, 4.0],   # w = .75 .5 .25 -> A 1.0, B .5 of 1.5
                  [0.0, 1.0, 2.0, 2.0],   # w = 1 .5 0 -> B 1.5 of 1.5
                  [1.0, 1.0, 1.0, 2.0],   # w = .5 .5 .5 -> thirds
                  [0.0, 0.0, 0.0, 0.0]])  # max 0 -> NaN weights -> NA
    assert target_classes(idx, d, groups, 0.5) == ["A", "B", "NA", "NA"]
    assert target_classes(idx, d, groups, 0.7) == ["NA", "B", "NA", "NA"]
    # two classes above a low threshold -> NA; one third is not above one third
    assert target_classes(idx, d, groups, 0.3) == ["NA", "B", "NA", "NA"]
    assert target_classes(idx, d, groups, 0.3333)[2] == "NA"
    assert target_classes(idx, d, groups, 0.5, target_subset=[1, 3]) == ["B", "NA"]
    # an exact half is not above one half
    assert target_classes(np.array([[0, 1, 2]]), np.array([[1.0, 1.0, 2.0]]), groups, 0.5) == ["NA"]


# ---- MiniBatchKMeans seed partitions (ann.py:328-346) -------------------------------
@pytest.mark.parametrize("case", ["blobs_k8", "small_blocks_k40", "init_subsample",
                                  "explicit_init_size", "k_equals_first_block",
                                  "outliers_cap", "two_clusters_big_init"])
def test_kmeans_oracle_matches_sklearn(case):
    """The restated MiniBatchKMeans reproduces scikit-learn's seeds, labels, counts and
    centers bit for bit, through the init subsample, empty-cluster reassignments and the
    0.5 * batch reassignment cap."""
    import warnings

    from sklearn.cluster import MiniBatchKMeans
    from kmeans_cases import CASES, blocks_of
    from oracle import kmeans as okm

    X, k, bs, block, init_size, first = CASES[case]()
    ref = MiniBatchKMeans(n_clusters=k, random_state=7, batch_size=bs, n_init=3,
                          init_size=init_size)
    mine = okm.MiniBatchKMeans(n_clusters=k, random_state=7, batch_size=bs,
                               init_size=init_size)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for b in blocks_of(X, block, first):
            ref.partial_fit(b)
            mine.partial_fit(b)
    assert np.array_equal(ref._counts, mine._counts)
    assert np.array_equal(ref.cluster_centers_, mine.cluster_centers_)
    assert np.array_equal(ref.predict(X), mine.predict(X))
    if case == "outliers_cap":
        sizes = [len(c) for (_, c, _) in mine.reassign_log]
        assert sizes and sizes[0] == 25  # 40 empty outlier centers, capped at 0.5 * 50


def test_kmeans_plusplus_oracle_matches_sklearn():
    from sklearn.cluster import kmeans_plusplus
    from kmeans_cases import blobs
    from oracle import kmeans as okm

    X = blobs(9, 700, 6, 12)
    _, want = kmeans_plusplus(X, 50, random_state=np.random.RandomState(3))
    _, got = okm.kmeans_plusplus(X, 50, okm.row_norms_sq(X), np.random.RandomState(3))
    assert np.array_equal(want, got)


def test_kmeans_update_is_sequential_chain():
    """c*w, + x in row order, * (1 / (w + n)): (0.1*3 + 0.2 + 0.7) * (1/5) differs from the
    plain mean (0.3 + 0.2 + 0.7) / 5 in the last bit, and an empty center is untouched."""
    from oracle import kmeans as okm

    X = np.array([[0.2], [0.7], [5.0]])
    C = np.array([[0.1], [9.0]])
    w = np.array([3.0, 4.0])
    okm.minibatch_update(X, C, w, np.array([0, 0, 0], dtype=np.int32))
    assert C[0, 0] == ((0.1 * 3.0 + 0.2) + 0.7 + 5.0) * (1.0 / 6.0)
    assert C[1, 0] == 9.0 and w.tolist() == [6.0, 4.0]


# ---- graph loading (graph_datastore.py:474-511, 1022-1075) ------------------------
def test_load_graph_oracle_hand_case():
    """Three cells, k = 2: 0->1 (0.5) and 1->0 (0.25) are mutual, 0->2 (0.8) is one-way,
    2->1 holds an explicit zero, 1->2 (1.0) meets 2->1's zero."""
    from oracle import graph as og

    edges = np.array([[0, 1], [0, 2], [1, 0], [1, 2], [2, 1], [2, 0]], dtype=np.uint64)
    w = np.array([0.5, 0.8, 0.25, 1.0, 0.0, 0.3])
    g = og.load_graph(edges, w, 3, 2)
    assert g.nnz == 6 and g[2, 1] == 0.0  # the explicit zero is kept
    t = og.load_graph(edges, w, 3, 2, symmetric=True).toarray()
    want = np.array([[0.0, 0.5 + 0.25 - 0.5 * 0.25, 0.8 + 0.3 - 0.8 * 0.3],
                     [0.5 + 0.25 - 0.5 * 0.25, 0.0, 1.0],
                     [0.8 + 0.3 - 0.8 * 0.3, 1.0, 0.0]])
    np.testing.assert_array_equal(t, want)
    assert (t == t.T).all()
    u = og.load_graph(edges, w, 3, 2, symmetric=True, upper_only=True)
    assert u.format == "coo" and (u.row <= u.col).all() and u.nnz == 3
    g1 = og.load_graph(edges, w, 3, 2, use_k=1)
    assert g1.nnz == 3 and g1[0, 1] == 0.5 and g1[1, 0] == 0.25 and g1[2, 1] == 0.0


# ---- CORAL (mapping_utils.py:15-95) ----------------------------------------------
def test_coral_oracle_pins():
    from oracle import coral as oc

    rng = np.random.default_rng(3)
    x = rng.normal(size=(400, 12)) @ rng.normal(size=(12, 12))
    y = rng.normal(size=(300, 12)) * np.linspace(0.5, 3, 12) + 2
    c = oc.cov_diaged(x)  # ddof 1, plus the identity
    m = x.mean(0)
    np.testing.assert_allclose(c, (x - m).T @ (x - m) / 399 + np.eye(12), rtol=1e-12,
                               atol=1e-12)
    s = (x - x.mean(0)) / x.std(0)
    t = (y - y.mean(0)) / y.std(0)
    A = oc.alignment(s, t)
    Cs, Ct = oc.cov_diaged(s), oc.cov_diaged(t)
    np.testing.assert_allclose(A.T @ Cs @ A, Ct, rtol=1e-9, atol=1e-9)  # re-colouring
    w, V = np.linalg.eigh(Cs)
    wt, Vt = np.linalg.eigh(Ct)
    np.testing.assert_allclose(A, (V * w ** -0.5) @ V.T @ (Vt * wt ** 0.5) @ Vt.T,
                               rtol=1e-9, atol=1e-10)
    data, A2 = oc.coral(x, x)  # same distribution: nothing to correct
    np.testing.assert_allclose(A2, np.eye(12), atol=1e-10)
    np.testing.assert_allclose(data, (x - x.mean(0)) / x.std(0), atol=1e-9)


def test_csr_stats_match_dense():
    """The sparse restatement used at full size equals the dense oracle (all-zero rows,
    all-zero and constant features included)."""
    from scarf_amd.synth import CSR, generate_csr, make_gene_model
    csr = generate_csr(700, make_gene_model(90, 3, 0.15, 3), 3)
    dense = csr.to_dense()
    dense[5] = 0
    dense[:, 7] = 0
    dense[:, 11] = 4
    c = CSR.from_dense(dense.astype(np.uint32))
    for sf in (1000.0, 1e4):
        x = norm.normed(dense, sf=sf)
        mu, sigma = norm.mu_sigma(x)
        m2, s2 = norm.csr_mu_sigma(c.vals, c.cols, c.rowptr, c.n_cols, sf=sf)
        np.testing.assert_allclose(m2, mu, rtol=1e-13, atol=1e-16)
        np.testing.assert_allclose(s2, sigma, rtol=1e-12, atol=1e-16)
        assert s2[7] == 1.0 and m2[7] == 0.0
        rows = np.array([0, 5, 699, 3])
        z = norm.csr_z_rows(c.vals, c.cols, c.rowptr, c.n_cols, rows, mu, sigma, sf=sf, clip=3.0)
        np.testing.assert_allclose(z, norm.transform_z_clipped(x[rows], mu, sigma, 3.0),
                                   rtol=1e-14, atol=1e-15)
    np.testing.assert_array_equal(norm.csr_libsize(c.vals, c.cols, c.rowptr, c.n_cols),
                                  norm.libsize(dense))


def test_c_stats_and_zgram_match_numpy():
    """The C restatement used at full size: libsize, mu, sigma and Z^T Z with the column
    sums equal the numpy oracle on the dense matrix (clip on and off, an all-zero row,
    an all-zero feature, unsorted columns in one row)."""
    from oracle import stats as ostats
    from scarf_amd.synth import CSR, generate_csr, make_gene_model
    csr = generate_csr(900, make_gene_model(70, 3, 0.2, 4), 4)
    dense = csr.to_dense()
    dense[3] = 0
    dense[:, 9] = 0
    c = CSR.from_dense(dense.astype(np.uint32))
    a, b = int(c.rowptr[10]), int(c.rowptr[11])
    c.vals[a:b], c.cols[a:b] = c.vals[a:b][::-1].copy(), c.cols[a:b][::-1].copy()
    for sf, clip in ((1000.0, None), (1e4, 2.5)):
        ls, mu, sg = ostats.csr_stats(c.vals, c.cols, c.rowptr, c.n_cols, sf=sf)
        x = norm.normed(dense, sf=sf)
        m0, s0 = norm.mu_sigma(x)
        np.testing.assert_array_equal(ls, norm.libsize(dense))
        np.testing.assert_allclose(mu, m0, rtol=1e-13, atol=1e-16)
        np.testing.assert_allclose(sg, s0, rtol=1e-12)
        G, cs = ostats.csr_zgram(c.vals, c.cols, c.rowptr, c.n_cols, ls, mu, sg, sf=sf, clip=clip)
        z = norm.transform_z_clipped(x, mu, sg, clip)
        scale = np.abs(z).T @ np.abs(z)
        assert (np.abs(G - z.T @ z) / np.maximum(scale, 1e-300)).max() < 1e-13
        assert not G[9].any() and not G[:, 9].any()  # the all-zero feature has z = 0
        np.testing.assert_allclose(cs, z.sum(0), atol=1e-10)
        w, Vt = ostats.exact_pca(G, cs, z.shape[0], 5)
        zc = z - z.mean(0)
        _, S, V2 = np.linalg.svd(zc, full_matrices=False)
        np.testing.assert_allclose(w, S[:5] ** 2, rtol=1e-10)
        V2 = ipca.svd_flip_rows(V2[:5])
        np.testing.assert_allclose(Vt, V2, atol=1e-8)
