"""GPU eigensolvers of the merge matrix (Chebyshev-filtered block subspace iteration, the
default, and thick-restart block Lanczos) and their kernels against LAPACK / numpy.

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


def _solver(name):
    from scarf_amd.eigen import top_eigh_chebyshev, top_eigh_device
    return {"lanczos": top_eigh_device, "chebyshev": top_eigh_chebyshev}[name]


@pytest.mark.parametrize("solver", ["chebyshev", "lanczos"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_top_eigh_device_vs_lapack(cuda, case, solver):
    import torch
    M, nt = CASES[case]()
    d = M.shape[0]
    say("eigen", solver, case, "d", d, "n_top", nt)
    w_ref, V_ref = sla.eigh(M, subset_by_index=[d - nt, d - 1])
    w_ref, V_ref = w_ref[::-1], V_ref[:, ::-1].T
    Md = torch.from_numpy(M).to("cuda")
    w, Vt, info = _solver(solver)(Md, nt)
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


@pytest.mark.parametrize("solver", ["chebyshev", "lanczos"])
def test_top_eigh_device_deterministic(cuda, solver):
    import torch
    M = torch.from_numpy(_gram(700, 4000, 9)).to("cuda")
    w1, V1, _ = _solver(solver)(M, 11)
    w2, V2, _ = _solver(solver)(M, 11)
    assert np.array_equal(w1, w2) and np.array_equal(V1, V2)


@pytest.mark.parametrize("d,ncols,ldy", [(100, 16, 16), (2000, 96, 96), (2000, 192, 192),
                                         (3000, 256, 256), (517, 48, 64), (2000, 160, 192)])
def test_symm_block_step(cuda, d, ncols, ldy):
    """sh_symm_block: Y2 = alpha M Y1 + beta Y1 + gamma Y0 against numpy, columns past
    ncols untouched; bit-identical on a repeat."""
    import torch
    from scarf_amd import _lib
    from scarf_amd.device import ptr, stream_handle
    rng = np.random.default_rng(d + ncols)
    A = rng.standard_normal((d, d))
    M = A + A.T
    Y1 = rng.standard_normal((d, ldy))
    Y0 = rng.standard_normal((d, ldy))
    al, be, ga = 0.7, -1.3, 0.25
    ref = al * (M @ Y1[:, :ncols]) + be * Y1[:, :ncols] + ga * Y0[:, :ncols]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    Md, Y1d, Y0d = t(M), t(Y1), t(Y0)
    out = []
    for _ in range(2):
        Y2d = torch.full((d, ldy), 7.0, dtype=torch.float64, device="cuda")
        ws = torch.empty(_lib.lib().sh_symm_block_workspace(d, ncols, ldy), dtype=torch.uint8,
                         device="cuda")
        _lib.call("sh_symm_block", ptr(Md), d, ptr(Y1d), ptr(Y0d), ncols, ldy, al, be, ga,
                  ptr(Y2d), ptr(ws), ws.numel(), stream_handle())
        out.append(Y2d.cpu().numpy())
    Y2 = out[0]
    np.testing.assert_allclose(Y2[:, :ncols], ref, rtol=1e-12,
                               atol=1e-13 * np.sqrt(d) * np.abs(M).max() * 4)
    assert np.all(Y2[:, ncols:] == 7.0)
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("d,na,nb,lda", [(2000, 94, 94, 96), (3000, 187, 187, 192),
                                         (33, 5, 7, 16), (1000, 130, 64, 160)])
def test_block_gram(cuda, d, na, nb, lda):
    import torch
    from scarf_amd import _lib
    from scarf_amd.device import ptr, stream_handle
    rng = np.random.default_rng(na + nb)
    A = rng.standard_normal((d, lda))
    B = rng.standard_normal((d, lda))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    Ad, Bd = t(A), t(B)
    C = torch.empty((na, nb), dtype=torch.float64, device="cuda")
    ws = torch.empty(_lib.lib().sh_block_gram_workspace(d, na, nb), dtype=torch.uint8,
                     device="cuda")
    _lib.call("sh_block_gram", ptr(Ad), lda, na, ptr(Bd), lda, nb, d, ptr(C), ptr(ws),
              ws.numel(), stream_handle())
    np.testing.assert_allclose(C.cpu().numpy(), A[:, :na].T @ B[:, :nb], rtol=1e-12,
                               atol=1e-13 * np.sqrt(d) * 4)


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
