"""GPU eigensolver (block Lanczos on the merge matrix) against LAPACK.

Eigenvalues agree to rounding; eigenvectors agree to within the solver's residual
tolerance divided by the relative eigenvalue gap (the cases keep that below 1e-7);
signs are aligned before comparing, as svd_flip does downstream.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import say

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _spiked_bulk(d, n_spikes, n_cells, seed):
    """Gram-like spectrum: a Marchenko-Pastur bulk for n_cells x d noise (beta-Laguerre
    bidiagonal model) plus n_spikes well separated group components."""
    rng = np.random.default_rng(seed)
    a = np.sqrt(rng.chisquare(n_cells - np.arange(d)))
    b = np.sqrt(rng.chisquare(d - 1 - np.arange(d - 1)))
    lam = np.sort(sla.svdvals(np.diag(a) + np.diag(b, -1)) ** 2)[::-1]
    spikes = lam[0] * np.geomspace(60, 1.3, n_spikes)
    lam = np.sort(np.concatenate([spikes, lam[:d - n_spikes]]))[::-1]
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    M = (Q * lam) @ Q.T
    return 0.5 * (M + M.T)


def _gram(d, n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)) * np.geomspace(3, 1, d)
    return X.T @ X


CASES = {
    "spiked_1m_2000": lambda: (_spiked_bulk(2000, 19, 1_000_000, 5), 51),
    "spiked_100k_1000": lambda: (_spiked_bulk(1000, 7, 100_000, 6), 31),
    "gram_800": lambda: (_gram(800, 5000, 7), 21),
    "gram_600_top1": lambda: (_gram(600, 3000, 8), 1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_top_eigh_device_vs_lapack(cuda, case):
    import torch
    from scarf_amd.eigen import top_eigh_device
    M, nt = CASES[case]()
    d = M.shape[0]
    say("eigen", case, "d", d, "n_top", nt)
    w_ref, V_ref = sla.eigh(M, subset_by_index=[d - nt, d - 1])
    w_ref, V_ref = w_ref[::-1], V_ref[:, ::-1].T
    Md = torch.from_numpy(M).to("cuda")
    w, Vt, info = top_eigh_device(Md, nt)
    say("  vectors", info.vectors, "cycles", info.cycles, "res", info.max_residual,
        "orth", info.orthogonality)
    np.testing.assert_allclose(w, w_ref, rtol=1e-12, atol=1e-13 * w_ref[0])
    s = np.sign(np.sum(Vt * V_ref, axis=1))
    gaps = np.abs(np.diff(np.concatenate([[np.inf], w_ref, [np.linalg.eigvalsh(M)[-nt - 1]]])))
    rel_gap = np.minimum(gaps[:-1], gaps[1:]) / w_ref[0]
    bound = np.maximum(1e-12 / rel_gap * 10, 1e-11)
    err = np.max(np.abs(s[:, None] * Vt - V_ref), axis=1)
    assert np.all(err <= bound), (err / bound).max()
    assert info.max_residual <= 1e-10 and info.orthogonality <= 1e-10


def test_top_eigh_device_deterministic(cuda):
    import torch
    from scarf_amd.eigen import top_eigh_device
    M = torch.from_numpy(_gram(700, 4000, 9)).to("cuda")
    w1, V1, _ = top_eigh_device(M, 11)
    w2, V2, _ = top_eigh_device(M, 11)
    assert np.array_equal(w1, w2) and np.array_equal(V1, V2)


def test_merge_matrix_matches_host(cuda):
    import torch
    from scarf_amd import _lib
    from scarf_amd.device import ptr, stream_handle
    rng = np.random.default_rng(10)
    d, nc, n = 300, 17, 250.0
    X = rng.standard_normal((400, d))
    G = X.T @ X
    bm = rng.standard_normal(d)
    Vt = np.linalg.qr(rng.standard_normal((d, nc)))[0].T
    s2 = rng.uniform(1, 100, nc)
    mc = rng.standard_normal(d)
    ref = (G - n * np.outer(bm, bm)) + (Vt.T * s2) @ Vt + np.outer(mc, mc)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    Gd, M = t(G), torch.empty((d, d), dtype=torch.float64, device="cuda")
    keep = [t(bm), t(Vt), t(s2), t(mc)]
    _lib.call("sh_merge_matrix", ptr(Gd), d, n, ptr(keep[0]), ptr(keep[1]), ptr(keep[2]), nc,
              ptr(keep[3]), ptr(M), stream_handle())
    Mh = M.cpu().numpy()
    assert np.array_equal(Mh, Mh.T)
    np.testing.assert_allclose(Mh, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("transA", [0, 1])
def test_dgemm(cuda, transA):
    import torch
    from scarf_amd import _lib
    from scarf_amd.device import ptr, stream_handle
    rng = np.random.default_rng(11 + transA)
    for m, n, k in ((51, 2000, 768), (70, 130, 3), (1, 1, 1), (64, 64, 64), (200, 65, 129)):
        A = rng.standard_normal((k, m) if transA else (m, k))
        B = rng.standard_normal((k, n))
        C0 = rng.standard_normal((m, n))
        ref = 0.5 * ((A.T if transA else A) @ B) + 2.0 * C0
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        Ad, Bd, Cd = t(A), t(B), t(C0)
        _lib.call("sh_dgemm", transA, m, n, k, 0.5, ptr(Ad), A.shape[1], ptr(Bd), n, 2.0,
                  ptr(Cd), n, stream_handle())
        np.testing.assert_allclose(Cd.cpu().numpy(), ref, rtol=1e-12,
                                   atol=1e-13 * np.sqrt(k) * 4)


def test_ipca_wide_device_eigensolver(cuda):
    """scikit-learn IncrementalPCA golden (600 genes, 4 partial fits) through the GPU
    eigensolver for every partial fit."""
    from scarf_amd.device import DeviceCSR
    from scarf_amd.normalize import NormParams, feature_stats
    from scarf_amd.pca import Projector, fit_pca
    from scarf_amd.synth import CSR
    g = np.load(os.path.join(GOLD, "ipca_wide.npz"))
    csr = CSR(g["vals"], g["cols"], g["rowptr"], int(g["n_cols"]))
    p = NormParams(size_factor=float(g["sf"]))
    dev = DeviceCSR.from_host(csr)
    chunk, dims = int(g["chunk"]), int(g["dims"])
    st = feature_stats(dev, chunk, p)
    fit = fit_pca(dev, st, p, dims, chunk, eig="device")
    say("  eig log", fit.state.eig_log)
    assert [b for b, _ in fit.state.eig_log] == ["device"] * len(fit.state.eig_log)
    proj = Projector.build(fit.loadings, st.mu.cpu().numpy(), st.sigma.cpu().numpy())
    E = proj.project(dev, p, libsize=st.libsize).cpu().numpy().astype(np.float64)
    Er = g["embedding"]
    for j in range(dims):
        s = np.sign(np.dot(E[:, j], Er[:, j])) or 1.0
        rel = np.linalg.norm(s * E[:, j] - Er[:, j]) / np.linalg.norm(Er[:, j])
        assert rel < 1e-4, (j, rel)
    np.testing.assert_allclose(fit.state.singular_values_, g["singular_values"], rtol=1e-9)
    np.testing.assert_allclose(fit.state.explained_variance_ratio_,
                               g["explained_variance_ratio"], rtol=1e-9)
    np.testing.assert_allclose(fit.state.mean_, g["ipca_mean"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(fit.state.var_, g["ipca_var"], rtol=1e-9)
