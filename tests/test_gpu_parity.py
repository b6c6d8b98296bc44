"""GPU parity: every HIP stage against the CPU oracle on the same seeded inputs."""
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
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_counts(n=3000, d=300, groups=4, seed=21, zero_rows=60):
    model = make_gene_model(d, groups, 0.1, seed)
    csr = generate_csr(n, model, seed)
    dense = csr.to_dense().astype(np.uint32)
    rng = np.random.default_rng(seed)
    dense[rng.choice(n, zero_rows, replace=False)] = 0
    return CSR.from_dense(dense), dense


@pytest.fixture(scope="module")
def data():
    return small_counts()


def test_norm_stats_parity(cuda, data):
    import torch
    from scarf_amd.device import DeviceCSR, col_map_tensor
    from scarf_amd.normalize import NormParams, feature_stats
    csr, dense = data
    feat = np.sort(np.random.default_rng(0).choice(dense.shape[1], 220, replace=False))
    for sf, chunk in ((1000.0, 700), (10000.0, 3000)):
        say("norm_stats sf", sf, "chunk", chunk)
        dev = DeviceCSR.from_host(csr)
        cm, d = col_map_tensor(csr.n_cols, feat)
        st = feature_stats(dev, chunk, NormParams(size_factor=sf), col_map=cm, d=d)
        torch.cuda.synchronize()
        x = onorm.normed(dense.astype(np.float64), sf=sf, feat_idx=feat)
        mu, sigma = onorm.mu_sigma(x)
        ls = onorm.libsize(dense.astype(np.float64), feat)
        np.testing.assert_array_equal(st.libsize.cpu().numpy(), ls)
        np.testing.assert_allclose(st.mu.cpu().numpy(), mu, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(st.sigma.cpu().numpy(), sigma, rtol=1e-12)


def test_norm_stats_deterministic(cuda, data):
    from scarf_amd.device import DeviceCSR
    from scarf_amd.normalize import NormParams, feature_stats
    csr, _ = data
    dev = DeviceCSR.from_host(csr)
    a = feature_stats(dev, 500, NormParams())
    b = feature_stats(dev, 500, NormParams())
    assert bool((a.mu == b.mu).all()) and bool((a.sigma == b.sigma).all())


@pytest.mark.parametrize("precision,method", [("fp64", "dense"), ("fp32", "dense"),
                                              ("fp64", "sparse")])
def test_gram_parity(cuda, data, precision, method):
    import torch
    from scarf_amd.device import DeviceCSR
    from scarf_amd.normalize import NormParams, feature_stats
    from scarf_amd.pca import gram_of_rows
    csr, dense = data
    dev = DeviceCSR.from_host(csr)
    p = NormParams()
    st = feature_stats(dev, 1000, p)
    rows = torch.arange(100, 2900, device="cuda", dtype=torch.int64)
    say("gram rows 100..2900 d", csr.n_cols, precision, method)
    G, cs, n = gram_of_rows(dev, st, p, rows, sub_rows=1000, precision=precision, method=method)
    x = onorm.normed(dense.astype(np.float64), sf=1000.0)
    mu, sigma = onorm.mu_sigma(x)
    z = ((x - mu) / sigma)[100:2900]
    Gr = z.T @ z
    # elementwise against the summation bound  c * u * sum_r |z_ri z_rj|
    scale = np.abs(z).T @ np.abs(z)
    err = (np.abs(G - Gr) / scale).max()
    say("gram max error / sum|z z|", err)
    assert err < (1e-12 if precision == "fp64" else 2e-5)  # FP64 bound ~ n u = 3e-13
    assert n == 2800
    np.testing.assert_allclose(cs, z.sum(0), atol=1e-9 if precision == "fp64" else 2e-3)


@pytest.mark.parametrize("name", ["ipca_small", "ipca_blocks", "ipca_wide"])
def test_ipca_golden_parity(cuda, name):
    """Loadings and PCA scores vs scikit-learn IncrementalPCA in scarf's order."""
    from scarf_amd.device import DeviceCSR
    from scarf_amd.normalize import NormParams, feature_stats
    from scarf_amd.pca import Projector, fit_pca
    g = np.load(os.path.join(GOLD, name + ".npz"))
    csr = CSR(g["vals"], g["cols"], g["rowptr"], int(g["n_cols"]))
    p = NormParams(size_factor=float(g["sf"]))
    dev = DeviceCSR.from_host(csr)
    chunk, dims = int(g["chunk"]), int(g["dims"])
    say("ipca golden", name)
    st = feature_stats(dev, chunk, p)
    np.testing.assert_allclose(st.mu.cpu().numpy(), g["mu"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(st.sigma.cpu().numpy(), g["sigma"], rtol=1e-12)
    fit = fit_pca(dev, st, p, dims, chunk)
    L = fit.loadings
    # scores: E = z @ loadings; compare per component after sign alignment
    proj = Projector.build(L, st.mu.cpu().numpy(), st.sigma.cpu().numpy())
    E = proj.project(dev, p, libsize=st.libsize).cpu().numpy().astype(np.float64)
    Er = g["embedding"]
    for j in range(dims):
        s = np.sign(np.dot(E[:, j], Er[:, j])) or 1.0
        rel = np.linalg.norm(s * E[:, j] - Er[:, j]) / np.linalg.norm(Er[:, j])
        assert rel < 1e-4, (name, j, rel)
    np.testing.assert_allclose(fit.state.explained_variance_ratio_, g["explained_variance_ratio"],
                               rtol=1e-4)


def test_project_parity(cuda, data):
    from scarf_amd.device import DeviceCSR
    from scarf_amd.normalize import NormParams, feature_stats
    from scarf_amd.pca import Projector
    csr, dense = data
    dev = DeviceCSR.from_host(csr)
    rng = np.random.default_rng(4)
    for D, clip in ((50, None), (100, 10.0), (13, None)):
        say("project D", D, "clip", clip)
        p = NormParams(clip=clip)
        st = feature_stats(dev, 1000, p)
        W = rng.standard_normal((csr.n_cols, D)) / np.sqrt(csr.n_cols)
        proj = Projector.build(W, st.mu.cpu().numpy(), st.sigma.cpu().numpy(), clip=clip)
        E = proj.project(dev, p, libsize=st.libsize).cpu().numpy()
        x = onorm.normed(dense.astype(np.float64), sf=1000.0)
        mu, sigma = onorm.mu_sigma(x)
        z = onorm.transform_z_clipped(x, mu, sigma, clip)
        Er = z @ W
        err = np.abs(E - Er).max() / np.abs(Er).max()
        say("project rel err", err)
        assert err < 1e-5


def _rand_embedding(n, D, seed, dup_rows=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((8, D)) * 3
    E = (c[rng.integers(0, 8, n)] + rng.standard_normal((n, D))).astype(np.float32)
    if dup_rows:
        E[rng.choice(n, dup_rows, replace=False)] = E[0]
    return E


@pytest.mark.parametrize("nq,nr,D,k,dup", [(700, 3000, 50, 11, 0), (300, 5000, 50, 21, 40),
                                           (50, 2000, 100, 11, 0), (1, 700, 13, 5, 0)])
def test_knn_fp32_exact_bitwise(cuda, nq, nr, D, k, dup):
    import torch
    from scarf_amd.knn import pack_refs, search
    R = _rand_embedding(nr, D, 5, dup)
    Q = np.concatenate([R[:nq // 2], _rand_embedding(nq - nq // 2, D, 6)])
    say("knn fp32 exact", nq, nr, D, k, dup)
    Rt, Qt = torch.from_numpy(R).cuda(), torch.from_numpy(Q).cuda()
    refs = pack_refs(Rt, "fp32")
    idx, dist, flags = search(Qt, Rt, refs, k, pool=max(32, k + 1))
    ri, rd = oknn.knn_f64(Q, R, k)
    np.testing.assert_array_equal(idx.cpu().numpy(), ri)
    np.testing.assert_array_equal(dist.cpu().numpy(), rd.astype(np.float32))


def test_knn_fp32_slices_and_self(cuda):
    import torch
    from scarf_amd.knn import pack_refs, search
    R = _rand_embedding(6000, 50, 7, dup_rows=30)
    Rt = torch.from_numpy(R).cuda()
    refs = pack_refs(Rt, "fp32")
    for S in (1, 3, 16):
        say("knn self-query slices", S)
        idx, dist, flags = search(Rt[:900], Rt, refs, 11, pool=32, self_offset=0, n_slices=S)
        fi, fd, nm, _ = oknn.self_query_knn(R, 11, query_rows=np.arange(900))
        np.testing.assert_array_equal(idx.cpu().numpy(), fi)
        np.testing.assert_array_equal(dist.cpu().numpy(), fd.astype(np.float32))
        got_mis = int(((flags.cpu().numpy() & 2) > 0).sum())
        # misses = rows whose first neighbour is not themselves (ann.py:34-35); the
        # flag marks those that had to drop their last column; compare the latter
        ki, _ = oknn.knn_f64(R[:900], R, 12)
        drop = sum(1 for i in range(900) if i not in ki[i])
        assert got_mis == drop


@pytest.mark.parametrize("n_slices", [1, 4])
def test_knn_fp32_uncertified_rows_fall_back(cuda, n_slices):
    """Near-tie data with a pool smaller than the tie group: the re-rank cannot certify
    those rows (flag 1), with one reference slice as with several, and their FP64
    brute-force recomputation equals the oracle."""
    import torch
    from scarf_amd.knn import pack_refs, search
    rng = np.random.default_rng(31)
    R = rng.standard_normal((3000, 24)).astype(np.float32)
    R[100:400] = R[100]  # 300 identical rows
    R[500:540] = R[500] + rng.standard_normal((40, 24)).astype(np.float32) * 1e-6
    Q = np.concatenate([R[[100, 250, 500, 520]], rng.standard_normal((60, 24)).astype(np.float32)])
    Rt, Qt = torch.from_numpy(R).cuda(), torch.from_numpy(Q).cuda()
    say("knn fp32 uncertified rows, slices", n_slices)
    refs = pack_refs(Rt, "fp32")
    idx, dist, flags = search(Qt, Rt, refs, 11, pool=16, n_slices=n_slices, order=False)
    f = flags.cpu().numpy()
    assert (f[:2] & 1).all(), f[:4]
    ri, rd = oknn.knn_f64(Q, R, 11)
    np.testing.assert_array_equal(idx.cpu().numpy(), ri)
    np.testing.assert_array_equal(dist.cpu().numpy(), rd.astype(np.float32))


def test_knn_bf16_recall(cuda):
    import torch
    from scarf_amd.knn import pack_refs, search
    R = _rand_embedding(20000, 50, 8)
    Q = R[:2000]
    Rt, Qt = torch.from_numpy(R).cuda(), torch.from_numpy(Q).cuda()
    say("knn bf16 pool 64")
    refs = pack_refs(Rt, "bf16")
    idx, dist, _ = search(Qt, Rt, refs, 11, pool=64)
    ri, rd = oknn.knn_f64(Q, R, 11)
    got = idx.cpu().numpy()
    recall = np.mean([len(set(got[i]) & set(ri[i])) / 11 for i in range(len(Q))])
    say("bf16 recall@11", recall)
    assert recall >= 0.999
    # returned distances are exact FP64 values of the returned neighbours
    pd = oknn.pair_dist_f64(Q, R, np.repeat(np.arange(len(Q)), 11), got.ravel())
    np.testing.assert_array_equal(dist.cpu().numpy().ravel(), pd.astype(np.float32))


@pytest.mark.parametrize("screen", ["fp32", "bf16", "fp8"])
def test_knn_locality_order_invariant(cuda, screen):
    """Screening the queries in locality order changes nothing: same indices, distances
    and flags as the natural order (self query with duplicates, and a plain query)."""
    import torch
    from scarf_amd.knn import locality_order, pack_refs, search
    R = _rand_embedding(30000, 50, 12, dup_rows=50)
    Rt = torch.from_numpy(R).cuda()
    refs = pack_refs(Rt, screen)
    for so, Q in ((0, Rt), (None, Rt[::3])):
        say("knn order", screen, "self" if so is not None else "plain", Q.shape[0])
        a = search(Q, Rt, refs, 21, pool=64, self_offset=so, order=False)
        b = search(Q, Rt, refs, 21, pool=64, self_offset=so, order=True)
        for x, y in zip(a, b):
            assert torch.equal(x, y)
    perm = locality_order(Rt).cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(R.shape[0]))


def test_knn_exact_kernel(cuda):
    import torch
    from scarf_amd.knn import exact_knn
    R = _rand_embedding(3000, 50, 9, dup_rows=20)
    Rt = torch.from_numpy(R).cuda()
    say("knn exact fp64 brute force")
    idx, dist, d64, _ = exact_knn(Rt[:500], Rt, 21)
    ri, rd = oknn.knn_f64(R[:500], R, 21)
    np.testing.assert_array_equal(idx.cpu().numpy(), ri)
    np.testing.assert_array_equal(d64.cpu().numpy(), rd)


def test_smooth_reference_fixture_gpu(cuda):
    import torch
    from scarf_amd.smooth import smooth_knn
    ki = np.load(os.path.join(GOLD, "knn_indices.npy")).astype(np.int64)
    kd = np.load(os.path.join(GOLD, "knn_distances.npy"))
    kw = np.load(os.path.join(GOLD, "knn_weights.npy"))
    say("smooth fixture")
    _, _, w = smooth_knn(torch.from_numpy(kd.astype(np.float32)).cuda(),
                         torch.from_numpy(ki).cuda(), 1000, 1.0, 1.5)
    w = w.cpu().numpy()
    assert np.abs(w - kw).max() < 1e-5


def test_smooth_chunked_vs_oracle(cuda):
    import torch
    from scarf_amd.smooth import smooth_knn
    R = _rand_embedding(2500, 20, 10, dup_rows=5)
    fi, fd, _, _ = oknn.self_query_knn(R, 11)
    say("smooth chunked 2500 rows chunk 300")
    _, wr = osmooth.smoothen_dists(fi, fd, 1.0, 1.5, 300)
    _, _, w = smooth_knn(torch.from_numpy(fd.astype(np.float32)).cuda(),
                         torch.from_numpy(fi).cuda(), 300, 1.0, 1.5)
    assert np.abs(w.cpu().numpy() - wr).max() < 1e-5


def test_mapping_scores_gpu(cuda):
    import torch
    from scarf_amd.mapping import mapping_scores
    ki = np.load(os.path.join(GOLD, "knn_indices.npy")).astype(np.int64)
    kd = np.load(os.path.join(GOLD, "knn_distances.npy"))
    fx = np.load(os.path.join(GOLD, "mapping_scores.npy"))
    n = ki.shape[0]
    idx = np.stack([np.arange(n), ki[:, 0], ki[:, 1]], 1)
    dist = np.stack([np.zeros(n), kd[:, 0], kd[:, 1]], 1).astype(np.float32)
    say("mapping scores fixture")
    (_, ms), = mapping_scores(torch.from_numpy(idx).cuda(), torch.from_numpy(dist).cuda(), n,
                              multiplier=3000.0)
    assert np.abs(ms - fx).max() < 1e-6  # float32 distances, as scarf stores hnswlib's
    groups = np.random.default_rng(0).integers(0, 3, n)
    got = mapping_scores(torch.from_numpy(idx).cuda(), torch.from_numpy(dist).cuda(), n,
                         groups=groups)
    want = omap.mapping_scores(idx, dist.astype(np.float64), n, target_groups=groups)
    for (g1, m1), (g2, m2) in zip(got, want):
        assert g1 == g2
        np.testing.assert_allclose(m1, m2, rtol=1e-11, atol=1e-12)


def _e4m3_table():
    vals = []
    for code in range(256):
        s, e, m = code >> 7, (code >> 3) & 15, code & 7
        if e == 15 and m == 7:
            continue  # NaN
        v = (m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
        vals.append((-v if s else v, code))
    return vals


def _e4m3_rne(x):
    """Nearest E4M3 code (ties to even mantissa), saturating at +-448."""
    tab = _e4m3_table()
    v = np.array([t[0] for t in tab])
    c = np.array([t[1] for t in tab])
    out = np.empty(x.shape, dtype=np.uint8)
    for i, xi in enumerate(x.ravel()):
        xi = min(max(xi, -448.0), 448.0)
        dist = np.abs(v - xi)
        best = np.where(dist == dist.min())[0]
        cand = c[best]
        if len(cand) > 1:
            even = cand[(cand & 1) == 0]
            cand = even if len(even) else cand
        cand = np.sort(cand)
        code = cand[0]
        if code & 0x7F == 0:  # rounds to zero: the sign is kept (-0 for negatives)
            code = 0x80 if xi < 0 else 0
        out.flat[i] = code
    return out


def test_fp8_pack_rounding(cuda):
    """The FP8 packer's E4M3 codes and E8M0 scales equal a round-to-nearest-even
    quantiser applied to x / 2^e."""
    import torch
    from scarf_amd.knn import pack
    rng = np.random.default_rng(11)
    E = (rng.standard_normal((40, 50)) * rng.lognormal(0, 2, (40, 1))).astype(np.float32)
    E[3] = 0.0
    E[5, :7] = [448.0, 1e-3, -2e-4, 0.0625, 3.0, -100.0, 1.0 / 1024]
    out, _, K = pack(torch.from_numpy(E).cuda(), "fp8", 0)
    b = out.cpu().numpy()
    for r in range(E.shape[0]):
        ex = int(b[r, K + 4:K + 8].view(np.int32)[0]) - 127
        want = _e4m3_rne(np.ldexp(E[r].astype(np.float64), -ex))
        assert np.abs(E[r]).max() == 0 or np.abs(np.ldexp(E[r], -ex)).max() <= 448
        np.testing.assert_array_equal(b[r, :50], want)
        assert b[r, K:K + 4].view(np.float32)[0] == np.float32((E[r].astype(np.float64) ** 2).sum())


def _fp8_model_recall(Q, R, pool, k=11):
    """Recall of an FP8 screen modelled on the CPU: rows quantised to E4M3 after a
    power-of-two scale (max |x| <= 448), keys ||r||^2 - 2 q.r in FP64, best `pool`
    by (key, index), exact re-rank."""
    tab = np.unique(np.array([t[0] for t in _e4m3_table()]))

    def q8(x):
        x = np.clip(x, -448, 448)
        i = np.clip(np.searchsorted(tab, x), 1, len(tab) - 1)
        lo, hi = tab[i - 1], tab[i]
        return np.where(np.abs(x - lo) <= np.abs(hi - x), lo, hi)

    def rows(X):
        ex = np.ceil(np.log2(np.maximum(np.abs(X).max(1), 1e-30) / 448))
        return q8(X / 2.0 ** ex[:, None]) * 2.0 ** ex[:, None]

    key = (R.astype(np.float64) ** 2).sum(1)[None, :] + rows(-2 * Q.astype(np.float64)) @ \
        rows(R.astype(np.float64)).T
    top = np.argsort(key, 1, kind="stable")[:, :pool]
    ri, _ = oknn.knn_f64(Q, R, k)
    rec = []
    for i in range(len(Q)):
        c = top[i]
        d = ((R[c].astype(np.float64) - Q[i]) ** 2).sum(1)
        rec.append(len(set(c[np.lexsort((c, d))][:k]) & set(ri[i])) / k)
    return float(np.mean(rec)), ri


@pytest.mark.parametrize("D,pool", [(100, 128), (50, 64)])
def test_knn_fp8_recall(cuda, D, pool):
    """FP8 screen recall equals the CPU model of the same quantisation (the kernel
    adds no error of its own); at D=100 with a 128 pool it exceeds 0.99.  With the
    margin test and the second screen of the flagged queries (the default for FP8) no
    query loses recall, and the refined rows are exact wherever the wider pool holds
    the true neighbours."""
    import torch
    from scarf_amd import _lib
    from scarf_amd.knn import pack_refs, search
    R = _rand_embedding(20000, D, 12)
    Q = R[:2000]
    Rt, Qt = torch.from_numpy(R).cuda(), torch.from_numpy(Q).cuda()
    say("knn fp8 D", D, "pool", pool)
    refs = pack_refs(Rt, "fp8")
    idx, dist, _ = search(Qt, Rt, refs, 11, pool=pool, refine=False)
    model, ri = _fp8_model_recall(Q, R, pool)
    got = idx.cpu().numpy()
    per_q = np.array([len(set(got[i]) & set(ri[i])) / 11 for i in range(len(Q))])
    recall = per_q.mean()
    say("fp8 recall@11", recall, "model", model)
    assert abs(recall - model) < 2e-3
    if D == 100:
        assert recall >= 0.99
    idx2, dist2, fl2 = search(Qt, Rt, refs, 11, pool=pool)
    got2 = idx2.cpu().numpy()
    per_q2 = np.array([len(set(got2[i]) & set(ri[i])) / 11 for i in range(len(Q))])
    thin = (fl2.cpu().numpy() & _lib.SH_FLAG_THIN) > 0
    say("refined", int(thin.sum()), "recall@11", per_q2.mean(),
        "misses left unflagged", int((per_q[~thin] < 1).sum()))
    # unflagged rows are the first screen's rows; flagged rows never lose recall
    assert np.array_equal(got2[~thin], got[~thin])
    assert np.all(per_q2[thin] >= per_q[thin])
    assert per_q2.mean() >= recall


@pytest.mark.parametrize("k,threshold", [(12, 0.5), (22, 0.5), (12, 0.3), (6, 0.9), (2, 0.5)])
def test_target_classes_gpu(cuda, k, threshold):
    """GPU label transfer equals the restated reference loop row for row (FP64 weights,
    numpy's summation order), including exact-threshold ties and zero-distance rows."""
    import torch
    from oracle.mapping import target_classes as ref_tc
    from scarf_amd.mapping import target_classes
    rng = np.random.default_rng(20 + k)
    n, n_ref, ncls = 3000, 500, 7
    groups = rng.integers(0, ncls, n_ref)
    idx = rng.integers(0, n_ref, (n, k))
    d = np.sort(rng.random((n, k)).astype(np.float32).astype(np.float64), axis=1)
    # ties: rows whose neighbours split evenly between two classes at equal distance
    d[:50] = 1.0
    d[:50, -1] = 2.0
    idx[:50, : k - 1] = np.where(np.arange(k - 1) % 2 == 0, 0, 1)
    groups[0], groups[1] = 0, 1
    d[50:60] = 0.0  # all-zero rows: NaN weights
    say("target classes k", k, "threshold", threshold)
    got = target_classes(torch.from_numpy(idx).cuda(), torch.from_numpy(d).cuda(),
                         torch.from_numpy(groups.astype(np.int32)), threshold).cpu().numpy()
    ref = ref_tc(idx, d, groups, threshold, na_val=-1)
    np.testing.assert_array_equal(got, np.array(ref))


def test_mapping_scores_beyond_one_fixed_point_batch(cuda):
    """More (query, neighbour) pairs than one 2^-40 fixed-point sum can hold at k = 11
    (1.6M target rows, 17.6M pairs): the entry point takes them in batches and adds the
    batch sums in FP64.  Against the oracle's per-group FP64 sums."""
    import torch
    from scarf_amd.mapping import mapping_scores
    rng = np.random.default_rng(44)
    nq, k, n_ref = 1_600_000, 11, 5000
    idx = rng.integers(0, n_ref, (nq, k)).astype(np.int64)
    dist = rng.uniform(0.0, 30.0, (nq, k)).astype(np.float32)
    groups = rng.integers(0, 3, nq)
    say("mapping scores", nq, "x", k, "pairs over", n_ref, "references, 3 groups")
    got = mapping_scores(torch.from_numpy(idx).cuda(), torch.from_numpy(dist).cuda(), n_ref,
                         groups=groups)
    w = 1.0 / (np.log1p(dist.astype(np.float64)) + 1.0)
    for g, ms in got:
        sel = groups == g
        s = np.bincount(idx[sel].ravel(), weights=w[sel].ravel(), minlength=n_ref)
        want = np.log1p(1000.0 * s / (sel.sum() * k))
        np.testing.assert_allclose(ms, want, rtol=1e-11)
    got_fw = mapping_scores(torch.from_numpy(idx).cuda(), torch.from_numpy(dist).cuda(), n_ref,
                            weighted=False, fixed_weight=0.1, log_transform=False)
    s = np.bincount(idx.ravel(), minlength=n_ref) * 0.1
    # each term is 0.1 rounded to a multiple of 2^-40: relative error <= 2^-41 / 0.1
    np.testing.assert_allclose(got_fw[0][1], 1000.0 * s / (nq * k), rtol=5e-12)
