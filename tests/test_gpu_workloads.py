"""GPU: the five workload configurations run at their full sizes.

* small (5,000 x 2,000, plain FP32 path) and 100k (100,000 x 2,000, IPCA 5 x 20k, BF16
  screen pool 64): mu/sigma and PCA against scikit-learn goldens written by
  tests/golden/make_config_golden.py on the same seeded counts, kNN and edge weights
  against the CPU oracle on the GPU's own embedding.
* 1M (1,000,000 x 2,000, k = 21, BF16 pool 64), serving (500,000-cell reference) and 4M
  (4,000,000 x 3,000, PCA 100, FP8 pool 128): the counts are drawn on the GPU, then
  checked through what does not need a dense copy -- mu/sigma and library sizes against
  the FP64 C restatement over the same CSR, the PCA against a host FP64 exact PCA of the
  same z (its Gram formed sparsely by the C oracle), PCA scores of sampled rows, the kNN
  of sampled queries against the FP64 brute force over every reference, the edge
  weights of one chunk against the oracle's smoothing, and (serving) each request's
  answer alone and in a mix.

References: ann.py:207-256 (PCA), ann.py:31-52 / 194-205 and knn_utils.py:23-76 (kNN),
knn_utils.py:89-159 (weights), mapping_datastore.py:31-285 (mapping).
"""
import os
import sys
import time

import numpy as np
import pytest

from oracle import knn as oknn
from oracle import norm as onorm
from oracle import smooth as osmooth
from oracle import stats as ostats

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _golden(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def _checksum(vals, cols, rowptr):
    return np.array([int(rowptr[-1]), int(vals.astype(np.uint64).sum()),
                     int(cols.astype(np.int64).sum()), int(rowptr.sum())], dtype=np.int64)


def _host_csr(dev):
    return (dev.vals.cpu().numpy().view(np.uint32), dev.cols.cpu().numpy(),
            dev.rowptr.cpu().numpy())  # (a no-op copy for host tensors)


def _check_components(L_gpu, L_ref, tol, what):
    """Per component after sign alignment: ||s l - l_ref|| / ||l_ref||."""
    for j in range(L_ref.shape[1]):
        s = np.sign(L_gpu[:, j] @ L_ref[:, j]) or 1.0
        rel = np.linalg.norm(s * L_gpu[:, j] - L_ref[:, j]) / np.linalg.norm(L_ref[:, j])
        assert rel < tol, (what, j, rel)


def _check_scores(E_rows, Er, tol, what):
    for j in range(Er.shape[1]):
        s = np.sign(E_rows[:, j] @ Er[:, j]) or 1.0
        rel = np.linalg.norm(s * E_rows[:, j] - Er[:, j]) / np.linalg.norm(Er[:, j])
        assert rel < tol, (what, j, rel)


def _check_knn_sample(res, E, rows, k, exact: bool, min_recall: float, what: str):
    """kNN of the sampled query rows (self queries: k + 1 then fix_knn_query) against the
    FP64 brute force over every reference row of the GPU's own embedding."""
    t0 = time.time()
    ri, rd = oknn.knn_f64(E[rows], E, k + 1)
    fi, fd, _ = oknn.fix_knn_query(ri, rd, rows)
    got = res.indices[rows].cpu().numpy()
    gd = res.distances[rows].cpu().numpy()
    if exact:
        np.testing.assert_array_equal(got, fi)
        np.testing.assert_array_equal(gd, fd.astype(np.float32))
        recall = 1.0
    else:
        recall = float(np.mean([len(set(got[i]) & set(fi[i])) / k for i in range(len(rows))]))
        # every returned distance is the exact FP64 distance of the returned neighbour
        pd = oknn.pair_dist_f64(E, E, np.repeat(rows, k), got.ravel())
        np.testing.assert_array_equal(gd.ravel(), pd.astype(np.float32))
        assert np.all(np.diff(gd, axis=1) >= 0)
    say(f"{what}: recall@{k} {recall:.6f} on {len(rows)} sampled queries "
        f"({time.time() - t0:.1f} s on the host)")
    assert recall >= min_recall, (what, recall)
    return fi


def _check_smooth_chunk(res, chunk, c, lc=1.0, bw=1.5):
    """Edge weights of chunk c against the oracle's smoothing of that chunk alone.  Rows
    are independent except for the chunk-local self index and the zero-weight rule,
    whose value (the global minimum over all chunks) is bounded by the chunk's own."""
    n, k = res.indices.shape
    a, b = c * chunk, min(n, (c + 1) * chunk)
    idx = res.indices[a:b].cpu().numpy()
    dist = res.distances[a:b].cpu().numpy()
    _, w_ref = osmooth.smoothen_dists(idx, dist, lc, bw, b - a)
    w = res.weights.view(-1)[a * k:b * k].cpu().numpy().astype(np.float64)
    _, _, vals = osmooth.compute_membership_strengths(
        idx, dist, *osmooth.smooth_knn_dist(dist.astype(np.float32), k, local_connectivity=lc,
                                            bandwidth=bw))
    nz = vals != 0
    np.testing.assert_allclose(w[nz], w_ref[nz], rtol=1e-5, atol=1e-6)
    if (~nz).any():
        assert np.all(w[~nz] > 0) and np.all(w[~nz] <= vals[nz].min() + 1e-6)


# ---- small: 5,000 x 2,000, plain FP32 path ---------------------------------------------
def test_config_small_full(cuda):
    import torch
    from scarf_amd.device import DeviceCSR
    from scarf_amd.graph import GraphParams, build_graph
    from scarf_amd.synth import generate_csr, make_gene_model
    g = _golden("config_small")
    n, d = int(g["n_cells"]), int(g["n_genes"])
    csr = generate_csr(n, make_gene_model(d, int(g["groups"]), float(g["density"]),
                                          int(g["seed"])), int(g["seed"]))
    np.testing.assert_array_equal(_checksum(csr.vals, csr.cols, csr.rowptr), g["checksum"])
    say("config small: build_graph 5000 x 2000, PCA 50 (IPCA, one chunk), FP32 exact k=11")
    p = GraphParams(dims=50, k=11, batch_size=int(g["chunk"]), size_factor=float(g["sf"]),
                    clip=float(g["clip"]), pca_mode="ipca", screen="fp32", pool=64)
    res = build_graph(DeviceCSR.from_host(csr), p)
    torch.cuda.synchronize()
    np.testing.assert_allclose(res.mu, g["mu"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(res.sigma, g["sigma"], rtol=1e-12)
    _check_components(res.loadings, g["loadings"], 1e-4, "small loadings")
    rows = g["score_rows"]
    E = res.embedding.cpu().numpy()
    _check_scores(E[rows].astype(np.float64), g["scores"], 1e-4, "small scores")
    np.testing.assert_allclose(res.pca.state.explained_variance_ratio_,
                               g["explained_variance_ratio"], rtol=1e-4)
    np.testing.assert_allclose(res.pca.state.singular_values_, g["singular_values"], rtol=1e-6)
    # every query, bit for bit against the FP64 brute force on the same embedding
    fi, fd, nm, rec = oknn.self_query_knn(E, 11)
    np.testing.assert_array_equal(res.indices.cpu().numpy(), fi)
    np.testing.assert_array_equal(res.distances.cpu().numpy(), fd.astype(np.float32))
    assert res.recall == pytest.approx(rec)
    _, w_ref = osmooth.smoothen_dists(fi, fd.astype(np.float32), 1.0, 1.5, int(g["chunk"]))
    np.testing.assert_allclose(res.weights.view(-1).cpu().numpy(), w_ref, rtol=1e-5, atol=1e-6)


# ---- 100k: IPCA 5 x 20k, BF16 screen pool 64 --------------------------------------------
def test_config_100k(cuda):
    import torch
    from scarf_amd import knn as gknn
    from scarf_amd.device import DeviceCSR
    from scarf_amd.graph import GraphParams, build_graph
    from scarf_amd.synth import generate_csr, make_gene_model
    g = _golden("config_100k")
    n, d = int(g["n_cells"]), int(g["n_genes"])
    t0 = time.time()
    csr = generate_csr(n, make_gene_model(d, int(g["groups"]), float(g["density"]),
                                          int(g["seed"])), int(g["seed"]))
    say(f"config 100k: counts drawn on the host in {time.time() - t0:.1f} s")
    np.testing.assert_array_equal(_checksum(csr.vals, csr.cols, csr.rowptr), g["checksum"])
    p = GraphParams(dims=50, k=11, batch_size=int(g["chunk"]), size_factor=float(g["sf"]),
                    clip=float(g["clip"]), pca_mode="ipca", screen="bf16", pool=64)
    say("config 100k: build_graph, IPCA 5 x 20000, BF16 screen pool 64, k=11")
    res = build_graph(DeviceCSR.from_host(csr), p)
    torch.cuda.synchronize()
    assert res.pca.mode == "ipca" and len(res.pca.state.eig_log) == 5
    # per-chunk FP64 sums merged with Chan's formula against numpy's two-pass mean/std of
    # the dense 100k x 2000 matrix: different summation orders over 20,000-row chunks
    np.testing.assert_allclose(res.mu, g["mu"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(res.sigma, g["sigma"], rtol=5e-12)
    _check_components(res.loadings, g["loadings"], 1e-4, "100k loadings")
    rows = g["score_rows"]
    E = res.embedding.cpu().numpy()
    _check_scores(E[rows].astype(np.float64), g["scores"], 1e-4, "100k scores")
    np.testing.assert_allclose(res.pca.state.explained_variance_ratio_,
                               g["explained_variance_ratio"], rtol=1e-4)
    q = np.sort(np.random.default_rng(7).choice(n, 2000, replace=False))
    _check_knn_sample(res, E, q, 11, exact=False, min_recall=0.999, what="100k bf16")
    # the FP32 screen on the same sampled queries is bit-exact
    Et = res.embedding
    qt = torch.from_numpy(q).cuda()
    i32, d32, _ = gknn.search(Et[qt].contiguous(), Et, gknn.pack_refs(Et, "fp32"), 11, pool=64,
                              self_idx=qt)
    ri, rd = oknn.knn_f64(E[q], E, 12)
    fi, fd, _ = oknn.fix_knn_query(ri, rd, q)
    np.testing.assert_array_equal(i32.cpu().numpy(), fi)
    np.testing.assert_array_equal(d32.cpu().numpy(), fd.astype(np.float32))
    _check_smooth_chunk(res, int(g["chunk"]), 3)


# ---- the GPU-drawn workloads: full-size property and sample checks ----------------------
def _gpu_workload(n, d, groups, density, seed):
    import torch
    from scarf_amd.device import DeviceCSR
    from scarf_amd.synth import generate_csr_gpu, make_gene_model
    t0 = time.time()
    model = make_gene_model(d, groups, density, seed=seed)
    vals, cols, rp = generate_csr_gpu(n, model, seed=seed,
                                      block_rows=min(50_000, max(1, 2_000_000_000 // (4 * d * 8))))
    csr = DeviceCSR.from_tensors(vals, cols, rp, d)
    del vals, cols, rp
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    say(f"drew {n} x {d} counts on the GPU (nnz {csr.nnz}) in {time.time() - t0:.1f} s")
    return model, csr


def _check_full_size(res, csr, sf, clip, chunk, k, exact_knn, min_recall, n_queries, what,
                     pca_tol=1e-4):
    """Stats, exact PCA, scores, sampled kNN and one chunk of weights of a GPU-drawn
    workload (exact PCA mode)."""
    n, d = csr.n_rows, csr.n_cols
    t0 = time.time()
    vals, cols, rp = _host_csr(csr)
    ls, mu, sg = ostats.csr_stats(vals, cols, rp, d, sf=sf)
    say(f"{what}: host FP64 stats in {time.time() - t0:.1f} s")
    np.testing.assert_array_equal(res.stats.libsize.cpu().numpy(), ls)
    np.testing.assert_allclose(res.mu, mu, rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(res.sigma, sg, rtol=1e-11)
    t0 = time.time()
    G, cs = ostats.csr_zgram(vals, cols, rp, d, ls, res.mu, res.sigma, sf=sf, clip=clip)
    nc = res.loadings.shape[1] + 1
    w, Vt = ostats.exact_pca(G, cs, n, nc)
    del G
    say(f"{what}: host FP64 exact PCA (sparse Gram + eigh) in {time.time() - t0:.1f} s")
    assert res.pca.mode == "exact"
    np.testing.assert_allclose(res.pca.state.singular_values_ ** 2, w, rtol=1e-9)
    _check_components(res.loadings, Vt[:-1].T, pca_tol, f"{what} loadings")
    rng = np.random.default_rng(7)
    rows = np.sort(rng.choice(n, 1000, replace=False))
    z = onorm.csr_z_rows(vals, cols, rp, d, rows, res.mu, res.sigma, sf=sf, clip=clip,
                         libsize=ls)
    E = res.embedding.cpu().numpy()
    _check_scores(E[rows].astype(np.float64), z @ res.loadings, 1e-5, f"{what} scores")
    del vals, cols, rp
    q = np.sort(rng.choice(n, n_queries, replace=False))
    _check_knn_sample(res, E, q, k, exact=exact_knn, min_recall=min_recall, what=what)
    _check_smooth_chunk(res, chunk, (n // chunk) // 2)
    return E


@pytest.mark.timeout(900)
def test_config_1m(cuda):
    import torch
    from scarf_amd.graph import GraphParams, build_graph
    _, csr = _gpu_workload(1_000_000, 2000, 20, 0.08, 1002)
    p = GraphParams(dims=50, k=21, batch_size=50_000, size_factor=1e4, clip=10.0,
                    pca_mode="exact", screen="bf16", pool=64)
    say("config 1m: build_graph 1M x 2000, exact PCA 50, BF16 pool 64, k=21")
    res = build_graph(csr, p)
    torch.cuda.synchronize()
    _check_full_size(res, csr, 1e4, 10.0, 50_000, 21, False, 0.999, 500, "1m")


def test_streamed_upload_matches_resident(cuda):
    """Counts streamed from pinned host memory in row chunks, statistics computed as each
    chunk lands: the resident copy is the host CSR and the statistics equal the one-launch
    pass over the resident matrix (a short last chunk included)."""
    import torch
    from scarf_amd.device import DeviceCSR
    from scarf_amd.graph import GraphParams, build_graph, build_graph_streamed
    from scarf_amd.normalize import HostCSR, NormParams, feature_stats, upload_with_stats
    from scarf_amd.synth import generate_csr, make_gene_model
    csr = generate_csr(9_100, make_gene_model(700, 4, 0.1, 61), 61)
    host = HostCSR.pinned_from(csr.vals.view(np.int32), csr.cols, csr.rowptr, csr.n_cols)
    p = NormParams(size_factor=1e4)
    say("streamed upload 9100 x 700 in chunks of 2000")
    dev, st = upload_with_stats(host, 2000, p)
    ref = DeviceCSR.from_host(csr)
    st0 = feature_stats(ref, 2000, p)
    torch.cuda.synchronize()
    assert torch.equal(dev.vals, ref.vals) and torch.equal(dev.cols, ref.cols)
    assert torch.equal(dev.rowptr, ref.rowptr)
    assert torch.equal(st.libsize, st0.libsize)
    np.testing.assert_allclose(st.mu.cpu().numpy(), st0.mu.cpu().numpy(), rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(st.sigma.cpu().numpy(), st0.sigma.cpu().numpy(), rtol=1e-13)
    gp = GraphParams(dims=20, k=11, batch_size=2000, size_factor=1e4, clip=10.0,
                     pca_mode="exact", screen="fp32")
    a = build_graph_streamed(host, gp)
    b = build_graph(ref, gp)
    _check_components(a.loadings, b.loadings, 1e-8, "streamed vs resident loadings")
    ai, bi = a.indices.cpu().numpy(), b.indices.cpu().numpy()
    assert np.mean(ai == bi) > 0.999


@pytest.mark.timeout(900)
def test_config_4m(cuda):
    """4M x 3000 streamed from pinned host memory in 100,000-cell chunks."""
    import torch
    from scarf_amd.graph import GraphParams, build_graph_streamed
    from scarf_amd.normalize import HostCSR
    _, dcsr = _gpu_workload(4_000_000, 3000, 40, 0.06, 1004)
    host = HostCSR.pinned_from(dcsr.vals, dcsr.cols, dcsr.rowptr, dcsr.n_cols)
    del dcsr
    torch.cuda.empty_cache()
    p = GraphParams(dims=100, k=11, batch_size=100_000, size_factor=1e4, clip=10.0,
                    pca_mode="exact", screen="fp8", pool=128)
    say("config 4m: streamed build 4M x 3000, exact PCA 100, FP8 pool 128, k=11")
    t0 = time.time()
    res = build_graph_streamed(host, p)
    torch.cuda.synchronize()
    say(f"config 4m: built in {time.time() - t0:.1f} s")
    from scarf_amd.device import DeviceCSR
    csr = DeviceCSR(host.vals, host.cols, host.rowptr, host.n_cols)  # host view for checks
    _check_full_size(res, csr, 1e4, 10.0, 100_000, 11, False, 0.995, 300, "4m")


@pytest.mark.timeout(900)
def test_config_serving_500k(cuda):
    """500,000-cell reference resident; requests of log-uniform sizes in [1, 50000] from the
    same group model with a 0.2 batch shift, in a shuffled gene order with 10 genes
    missing.  Each request's answer is bit-identical alone and in a concurrent mix; the
    projection of sampled query cells matches the FP64 host value and their neighbours
    match the FP64 brute force over the whole reference."""
    import torch
    from scarf_amd.device import DeviceCSR
    from scarf_amd.graph import GraphParams, build_graph
    from scarf_amd.mapping import align_col_map
    from scarf_amd.pca import Projector
    from scarf_amd.device import col_map_from_positions
    from scarf_amd.serving import MappingServer
    from scarf_amd.synth import CSR, generate_csr_gpu
    d = 2000
    model, ref = _gpu_workload(500_000, d, 20, 0.08, 1003)
    p = GraphParams(dims=50, k=11, batch_size=50_000, size_factor=1e4, clip=10.0,
                    pca_mode="exact", screen="bf16", pool=64)
    say("serving 500k: reference build_graph")
    res = build_graph(ref, p)
    torch.cuda.synchronize()
    del ref
    rng = np.random.default_rng(1003)
    qcols = rng.permutation(d)[: d - 10]
    src_ids = [f"g{i}" for i in range(d)]
    q_ids = [f"g{i}" for i in qcols]
    qv, qc, qr = generate_csr_gpu(120_000, model, seed=1503, batch_shift_sigma=0.2)
    pool = CSR(qv.cpu().numpy().view(np.uint32), qc.cpu().numpy(), qr.cpu().numpy(), d)
    del qv, qc, qr
    dense_cols = np.full(d, -1, np.int64)
    dense_cols[qcols] = np.arange(d - 10)

    def request(a, b):  # rows [a, b) of the pool in the query's gene order
        sub = pool.row_slice(a, b)
        keep = dense_cols[sub.cols] >= 0
        nc = dense_cols[sub.cols][keep].astype(np.int32)
        rows = np.repeat(np.arange(b - a), np.diff(sub.rowptr))[keep]
        order = np.lexsort((nc, rows))
        rp = np.zeros(b - a + 1, np.int64)
        np.cumsum(np.bincount(rows, minlength=b - a), out=rp[1:])
        return CSR(sub.vals[keep][order], nc[order], rp, d - 10)

    sizes = np.exp(rng.uniform(0, np.log(50_000), 8)).astype(np.int64).clip(1, 50_000)
    sizes[0], sizes[1] = 1, 50_000
    starts = rng.integers(0, 120_000 - 50_000, len(sizes))
    reqs = [request(int(s), int(s + z)) for s, z in zip(starts, sizes)]
    say("serving 500k: request sizes", sizes.tolist())
    index = res.index
    kw = dict(norm=p.norm(), k=11, mem_budget=8 << 30, chunk_rows=16384, n_streams=4)
    with MappingServer(res.loadings, res.mu, res.sigma, index, src_ids, q_ids, **kw) as srv:
        alone = [srv.submit(r).result(timeout=300) for r in reqs]
    with MappingServer(res.loadings, res.mu, res.sigma, index, src_ids, q_ids, **kw) as srv:
        futs = [srv.submit(r) for r in reqs]
        mixed = [f.result(timeout=300) for f in futs]
    for a, m, z in zip(alone, mixed, sizes):
        assert a.error is None and m.error is None
        assert a.indices.shape == (z, 11)
        np.testing.assert_array_equal(a.indices, m.indices)
        np.testing.assert_array_equal(a.distances, m.distances)
        np.testing.assert_array_equal(a.scores, m.scores)
    say("serving 500k: alone == mixed for", len(reqs), "requests")
    # sampled query cells of the largest request: projection and neighbours
    big = reqs[1]
    positions, missing = align_col_map(src_ids, q_ids)
    proj = Projector.build(res.loadings, res.mu, res.sigma, clip=10.0, fill_missing=missing)
    dev = DeviceCSR.from_host(big)
    Eq = proj.project(dev, p.norm(), col_map=col_map_from_positions(len(q_ids), positions))
    Eq = Eq.cpu().numpy()
    rows = np.sort(rng.choice(big.n_rows, 300, replace=False))
    # host FP64: libsize over the present features, x, missing features = 1.0 (align_features)
    xa = np.ones((len(rows), d))
    for t, r in enumerate(rows):
        a, b = int(big.rowptr[r]), int(big.rowptr[r + 1])
        c = big.vals[a:b].astype(np.float64)
        xr = np.zeros(d)
        xr[qcols] = 0.0
        xr[qcols[big.cols[a:b]]] = np.log1p(1e4 * c / (c.sum() or 1.0))
        xa[t] = np.where(missing, 1.0, xr)
    zq = np.clip((xa - res.mu) / res.sigma, -10.0, 10.0)
    _check_scores(Eq[rows].astype(np.float64), zq @ res.loadings, 1e-5, "serving projection")
    Er = res.embedding.cpu().numpy()
    ri, rd = oknn.knn_f64(Eq[rows], Er, 11)
    got = alone[1].indices[rows].astype(np.int64)
    recall = float(np.mean([len(set(got[i]) & set(ri[i])) / 11 for i in range(len(rows))]))
    say(f"serving 500k: recall@11 {recall:.6f} on {len(rows)} sampled query cells")
    assert recall >= 0.999
    pd = oknn.pair_dist_f64(Eq, Er, np.repeat(rows, 11), got.ravel())
    np.testing.assert_array_equal(alone[1].distances[rows].ravel(), pd.astype(np.float32))
