"""GPU: the drop-in layer (make_graph / AnnStream / knn_utils / run_mapping / GpuIndex)
chained stage by stage against the CPU oracle on the same inputs."""
import os
import sys

import numpy as np
import pytest

from oracle import ipca as oipca
from oracle import knn as oknn
from oracle import mapping as omap
from oracle import norm as onorm
from oracle import smooth as osmooth
from scarf_amd.synth import CSR, generate_csr, make_gene_model

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ref_data():
    model = make_gene_model(300, 4, 0.12, 31)
    csr = generate_csr(2400, model, 31)
    return model, csr


def _sk_loadings(x, feat_rows_blocks, dims, mu, sigma):
    from sklearn.decomposition import IncrementalPCA
    L, _ = oipca.fit_pca_scarf(feat_rows_blocks, dims, mu, sigma,
                               fitter=IncrementalPCA(n_components=dims + 1))
    return L


def test_make_graph_chain(cuda, ref_data):
    from scarf_amd.datastore import make_graph
    _, csr = ref_data
    feat = np.sort(np.random.default_rng(1).choice(300, 180, replace=False))
    say("make_graph chain fp32 screen")
    store = {}
    ann = make_graph(csr, feat, dims=20, k=11, batch_size=500, screen="fp32", store=store,
                     return_ann_object=True)
    dense = csr.to_dense()
    x = onorm.normed(dense, sf=1000.0, feat_idx=feat)
    mu, sigma = onorm.mu_sigma(x)
    np.testing.assert_allclose(ann.mu, mu, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(ann.sigma, sigma, rtol=1e-12)
    blocks = [x[s:s + 500] for s in range(0, x.shape[0], 500)]
    L = _sk_loadings(x, blocks, 20, mu, sigma)
    Er = ((x - mu) / sigma) @ L
    E = ann.annIdx.embedding().cpu().numpy().astype(np.float64)
    for j in range(20):
        s = np.sign(E[:, j] @ Er[:, j])
        assert np.linalg.norm(s * E[:, j] - Er[:, j]) / np.linalg.norm(Er[:, j]) < 1e-4
    np.testing.assert_allclose(ann._pca.explained_variance_ratio_[:20],
                               oipca.fit_pca_scarf(blocks, 20, mu, sigma)[1].explained_variance_ratio_[:20],
                               rtol=1e-6)
    # kNN: oracle on the GPU's own embedding -> bit-exact
    fi, fd, nm, rec = oknn.self_query_knn(E.astype(np.float32), 11)
    knn_loc = store["latest_knn"]
    np.testing.assert_array_equal(store[knn_loc + "/indices"], fi.astype(np.uint64))
    np.testing.assert_array_equal(store[knn_loc + "/distances"],
                                  fd.astype(np.float32).astype(np.float64))
    assert store[knn_loc + "/recall"] == pytest.approx(rec)
    # seed partitions: the oracle's MiniBatchKMeans on the same embedding, block by block
    # (n_centroids 1000 -> 500 clusters, the batch size; graph_datastore.py:744, 958-975)
    from oracle import kmeans as okm
    kloc = [v for key, v in store.items() if key.endswith("/latest_kmeans")][0]
    assert kloc.endswith("/kmeans__1000__4466")
    km = okm.MiniBatchKMeans(n_clusters=500, random_state=4466, batch_size=500)
    for s in range(0, E.shape[0], 500):
        km.partial_fit(E[s:s + 500])
    np.testing.assert_array_equal(store[kloc + "/cluster_centers"], km.cluster_centers_)
    np.testing.assert_array_equal(store[kloc + "/cluster_labels"],
                                  km.predict(E).astype(np.float64))
    # weights: oracle smoothing on the same kNN table
    e_ref, w_ref = osmooth.smoothen_dists(fi, fd.astype(np.float32), 1.0, 1.5, 500)
    g = store["latest_graph"]
    np.testing.assert_array_equal(store[g + "/edges"], e_ref)
    assert np.abs(store[g + "/weights"] - w_ref).max() < 1e-5


def test_make_graph_bf16_recall(cuda, ref_data):
    from scarf_amd.datastore import make_graph
    _, csr = ref_data
    say("make_graph bf16 screen")
    ann = make_graph(csr, None, dims=30, k=21, batch_size=800, screen="bf16",
                     return_ann_object=True)
    E = ann.annIdx.embedding().cpu().numpy()
    fi, _, _, _ = oknn.self_query_knn(E, 21)
    st = {}
    from scarf_amd.knn_utils import self_query_knn
    self_query_knn(ann, st, 800)
    got = st["indices"].astype(np.int64)
    rec = np.mean([len(set(got[i]) & set(fi[i])) / 21 for i in range(len(fi))])
    say("bf16 self recall@21", rec)
    assert rec >= 0.999


def test_run_mapping_and_scores(cuda, ref_data):
    from scarf_amd.datastore import get_mapping_score, make_graph, run_mapping
    model, csr = ref_data
    feat = np.arange(0, 300, 2)  # 150 source features
    src_ids = [f"g{i}" for i in feat]
    ann = make_graph(csr, feat, dims=15, k=11, batch_size=600, screen="fp32",
                     return_ann_object=True)
    # target: new cells from the same model, columns shuffled, 10 source genes absent,
    # 20 extra genes the reference does not use
    tgt = generate_csr(700, model, 77, batch_shift_sigma=0.2)
    tdense = tgt.to_dense()
    rng = np.random.default_rng(3)
    keep = np.setdiff1d(np.arange(300), feat[rng.choice(150, 10, replace=False)])
    extra = np.setdiff1d(np.arange(300), feat)[:20]
    cols = np.concatenate([np.intersect1d(keep, feat), extra])
    rng.shuffle(cols)
    tcounts = tdense[:, cols].astype(np.uint32)
    tids = [f"g{i}" for i in cols]
    say("run_mapping 700 target cells")
    store = run_mapping(ann, CSR.from_dense(tcounts), tids, src_ids, "t1", save_k=5)
    # oracle: align, z with reference mu/sigma, project with the GPU loadings
    t_idx = omap.order_features(src_ids, tids)
    xa = omap.align_normed(tcounts, t_idx, sf=1000.0)
    Et = ((xa - ann.mu) / ann.sigma) @ ann.loadings
    Ef = ann.annIdx.embedding().cpu().numpy()
    # the GPU's target embedding (FP32 projection) against the oracle's FP64 one
    from scarf_amd.datastore import project_target
    Eg = project_target(ann, CSR.from_dense(tcounts), tids, src_ids).cpu().numpy()
    assert np.abs(Eg - Et).max() / np.abs(Et).max() < 1e-5
    # neighbours: bit for bit against the FP64 brute force on that same embedding
    i_got = store["RNA/projections/t1/indices"].astype(np.int64)
    ri, rd = oknn.knn_f64(Eg, Ef, 5)
    np.testing.assert_array_equal(i_got, ri)
    np.testing.assert_array_equal(store["RNA/projections/t1/distances"],
                                  rd.astype(np.float32).astype(np.float64))
    # scores: oracle formula on the stored table
    (_, ms), = list(get_mapping_score(store, "t1"))
    (_, msr), = omap.mapping_scores(i_got, store["RNA/projections/t1/distances"], ann.nCells)
    np.testing.assert_allclose(ms, msr, rtol=1e-10, atol=1e-12)
    # label transfer from the same table: reference classes from the first PC's sign and
    # magnitude (4 classes), GPU vote vs the restated reference loop
    from scarf_amd.datastore import get_target_classes
    e0 = Ef[:, 0]
    labels = np.where(e0 > np.median(e0), "hi", "lo").astype(object)
    labels[np.abs(e0) < np.percentile(np.abs(e0), 10)] = "mid"
    for thr in (0.5, 0.8):
        say("get_target_classes threshold", thr)
        pred = get_target_classes(store, "t1", labels, threshold_fraction=thr)
        ref = omap.target_classes(i_got, store["RNA/projections/t1/distances"], labels, thr)
        assert list(pred) == ref
    sub = get_target_classes(store, "t1", labels, target_subset=[5, 1, 600])
    assert list(sub) == omap.target_classes(i_got, store["RNA/projections/t1/distances"],
                                            labels, 0.5, target_subset=[5, 1, 600])


def test_target_projection_parity(cuda, ref_data):
    """Target embedding itself (alignment + missing = 1.0 + subset library size)."""
    import torch
    from scarf_amd.device import DeviceCSR, col_map_from_positions
    from scarf_amd.mapping import align_col_map
    from scarf_amd.normalize import NormParams
    from scarf_amd.pca import Projector
    model, _ = ref_data
    tgt = generate_csr(300, model, 78).to_dense()
    src = [f"g{i}" for i in range(0, 300, 3)]
    cols = np.arange(1, 300, 2)
    tids = [f"g{i}" for i in cols]
    tcounts = tgt[:, cols].astype(np.uint32)
    rng = np.random.default_rng(5)
    L = rng.standard_normal((len(src), 12))
    mu = rng.uniform(0, 1, len(src))
    sigma = rng.uniform(0.5, 2, len(src))
    pos, missing = align_col_map(src, tids)
    proj = Projector.build(L, mu, sigma, fill_missing=missing)
    dcsr = DeviceCSR.from_host(CSR.from_dense(tcounts))
    E = proj.project(dcsr, NormParams(), libsize=None,
                     col_map=col_map_from_positions(dcsr.n_cols, pos)).cpu().numpy()
    xa = omap.align_normed(tcounts, omap.order_features(src, tids), sf=1000.0)
    Er = ((xa - mu) / sigma) @ L
    assert np.abs(E - Er).max() / np.abs(Er).max() < 1e-5


def test_gpu_index_api(cuda, tmp_path):
    from scarf_amd.knn import GpuIndex
    rng = np.random.default_rng(9)
    X = rng.standard_normal((3000, 24)).astype(np.float32)
    idx = GpuIndex(space="l2", dim=24, screen="fp32")
    idx.init_index(max_elements=3000, ef_construction=50, M=48, random_seed=4466)
    idx.set_ef(50)
    idx.set_num_threads(1)
    for s in range(0, 3000, 700):
        idx.add_items(X[s:s + 700])
    assert idx.get_current_count() == 3000
    lab, dist = idx.knn_query(X[:100], k=7)
    assert lab.dtype == np.uint64 and dist.dtype == np.float32
    ri, rd = oknn.knn_f64(X[:100], X, 7)
    np.testing.assert_array_equal(lab.astype(np.int64), ri)
    p = str(tmp_path / "ann_idx")
    idx.save_index(p)
    j = GpuIndex(space="l2", dim=24)
    j.load_index(p)
    lab2, dist2 = j.knn_query(X[:100], k=7)
    np.testing.assert_array_equal(lab, lab2)
    with pytest.raises(RuntimeError):
        idx.add_items(X[:10])  # exceeds max_elements
    with pytest.raises(RuntimeError):
        idx.knn_query(X[:1], k=5000)


def test_fix_knn_query_gpu(cuda):
    from scarf_amd.ann import fix_knn_query
    rng = np.random.default_rng(2)
    n, k1 = 400, 12
    ind = rng.integers(0, 1000, (n, k1)).astype(np.uint64)
    ref = rng.integers(0, 1000, n).astype(np.int64)
    where = rng.integers(-1, k1, n)
    for i, w in enumerate(where):
        ind[i][ind[i] == ref[i]] = 5000 + i  # make self absent
        if w >= 0:
            ind[i, w] = ref[i]
    dist = rng.random((n, k1)).astype(np.float32)
    a, b, c = fix_knn_query(ind, dist, ref)
    ra, rb, rc = oknn.fix_knn_query(ind.astype(np.int64), dist, ref)
    np.testing.assert_array_equal(a.astype(np.int64), ra)
    np.testing.assert_array_equal(b, rb)
    assert c == rc


def test_transform_ann_noncontiguous_self(cuda, ref_data):
    from scarf_amd.datastore import make_graph
    _, csr = ref_data
    ann = make_graph(csr, None, dims=10, k=8, batch_size=1000, screen="fp32",
                     return_ann_object=True)
    E = ann.annIdx.embedding()
    rows = np.array([5, 900, 17, 2399, 1200])
    i, d, nm = ann.transform_ann(E[rows].cpu().numpy(), k=8, self_indices=rows)
    ki, kd = oknn.knn_f64(E[rows].cpu().numpy(), E.cpu().numpy(), 9)
    fi, fd, fnm = oknn.fix_knn_query(ki, kd, rows)
    np.testing.assert_array_equal(i.astype(np.int64), fi)
    assert nm == fnm


@pytest.mark.parametrize("clip", [None, 3.0])
def test_dense_block_reducer_mapping_loop(cuda, ref_data, clip):
    """scarf's own mapping loop on dense numpy blocks of normalised values
    (mapping_datastore.py:198-205): ki, kd = ann.transform_ann(ann.reducer(a), k=save_k).
    transform_z equals numpy's (a - mu) / sigma bit for bit; the reducer is the FP64
    product within rounding; the neighbours equal the FP64 brute force on that embedding."""
    import torch
    from scarf_amd.datastore import make_graph
    model, csr = ref_data
    feat = np.arange(0, 300, 2)
    ann = make_graph(csr, feat, dims=15, k=11, batch_size=600, screen="fp32", clip=clip,
                     return_ann_object=True)
    Ef = ann.annIdx.embedding().cpu().numpy()
    tgt = generate_csr(700, model, 79, batch_shift_sigma=0.2).to_dense()[:, feat]
    x = onorm.normed(tgt, sf=1000.0)
    say("dense-block mapping loop, clip", clip)
    for s in range(0, 700, 250):
        a = x[s:s + 250]
        z_ref = (a - ann.mu) / ann.sigma
        if clip:
            z_ref = np.clip(z_ref, -clip, clip)
        z = ann.transform_z(a)
        assert isinstance(z, np.ndarray) and z.dtype == np.float64
        np.testing.assert_array_equal(z, z_ref)
        E = ann.reducer(a)
        assert isinstance(E, np.ndarray) and E.dtype == np.float64 and E.shape == (len(a), 15)
        Er = z_ref @ ann.loadings
        scale = np.abs(z_ref) @ np.abs(ann.loadings)
        assert (np.abs(E - Er) / scale).max() < 1e-13
        ki, kd = ann.transform_ann(E, k=5)
        ri, rd = oknn.knn_f64(E.astype(np.float32), Ef, 5)
        np.testing.assert_array_equal(ki.astype(np.int64), ri)
        np.testing.assert_array_equal(kd, rd.astype(np.float32))
    # FP32 blocks and device tensors take the same kernel
    a = x[:100]
    np.testing.assert_array_equal(ann.reducer(a.astype(np.float32)),
                                  ann.reducer(a.astype(np.float32).astype(np.float64)))
    Et = ann.reducer(torch.from_numpy(a).cuda())
    assert isinstance(Et, torch.Tensor) and Et.dtype == torch.float64
    np.testing.assert_array_equal(Et.cpu().numpy(), ann.reducer(a))
    np.testing.assert_array_equal(ann.transform_z(torch.from_numpy(a).cuda()).cpu().numpy(),
                                  ann.transform_z(a))
    with pytest.raises(ValueError):
        ann.reducer(x[:10, :-1])
    # mapping replaces mu / sigma (ref_mu / ref_sigma False, mapping_datastore.py:176-190)
    ann.mu = x.mean(0)
    ann.sigma = onorm.clean_array(x.std(0), 1)
    z_ref = (a - ann.mu) / ann.sigma
    if clip:
        z_ref = np.clip(z_ref, -clip, clip)
    np.testing.assert_array_equal(ann.transform_z(a), z_ref)


def test_dense_block_reducer_without_scaling(cuda, ref_data):
    """feat_scaling=False (disable_scaling, ann.py:119-127): reducer(a) = a @ loadings."""
    from scarf_amd.datastore import make_graph
    model, csr = ref_data
    ann = make_graph(csr, None, dims=10, k=7, batch_size=800, screen="fp32",
                     feat_scaling=False, return_ann_object=True)
    a = onorm.normed(generate_csr(120, model, 80).to_dense(), sf=1000.0)
    E = ann.reducer(a)
    Er = a @ ann.loadings
    assert (np.abs(E - Er) / (np.abs(a) @ np.abs(ann.loadings))).max() < 1e-13


def test_batch_size_and_k_clamps(cuda):
    """ann.py:85-86 (k >= nCells -> nCells - 1) and _handle_batch_size, ann.py:173-185
    (dims >= batch_size -> batch_size - 1, nClusters > batch_size -> batch_size), with the
    clamped graph checked against the oracle."""
    from scarf_amd.datastore import make_graph
    model = make_gene_model(60, 2, 0.3, 51)
    csr = generate_csr(40, model, 51)
    say("clamps: 40 cells, k=100, dims=30, batch_size=20")
    store = {}
    ann = make_graph(csr, None, dims=30, k=100, batch_size=20, screen="fp32", store=store,
                     return_ann_object=True)
    assert ann.k == 39
    assert ann.dims == 19 and ann.loadings.shape == (60, 19)
    assert ann.nClusters == 20
    x = onorm.normed(csr.to_dense(), sf=1000.0)
    mu, sigma = onorm.mu_sigma(x)
    blocks = [x[s:s + 20] for s in range(0, 40, 20)]
    L = _sk_loadings(x, blocks, 19, mu, sigma)
    E = ann.annIdx.embedding().cpu().numpy().astype(np.float64)
    Er = ((x - mu) / sigma) @ L
    for j in range(19):
        sg = np.sign(E[:, j] @ Er[:, j])
        assert np.linalg.norm(sg * E[:, j] - Er[:, j]) / np.linalg.norm(Er[:, j]) < 1e-4
    fi, fd, _, _ = oknn.self_query_knn(E.astype(np.float32), 39)
    knn_loc = store["latest_knn"]
    np.testing.assert_array_equal(store[knn_loc + "/indices"], fi.astype(np.uint64))
    # dims larger than the cell count is clamped to it (ann.py:174-175); fitting
    # dims + 1 components on 15 cells then fails, as IncrementalPCA's does
    small = generate_csr(15, model, 52)
    with pytest.raises(ValueError):
        make_graph(small, None, dims=30, k=5, batch_size=1000, screen="fp32",
                   return_ann_object=True)

