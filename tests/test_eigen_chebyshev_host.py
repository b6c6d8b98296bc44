"""Host loop of the Chebyshev-filtered block subspace eigensolver (eigen.chebyshev_subspace).

The GPU backend (_DeviceBlock: sh_symm_block, sh_block_gram, sh_dgemm, sh_block_residual)
is restated here with numpy in the same conventions (the active block as the first nb
columns of a d x ld array, locked vectors as rows deflated out of a working copy of M),
so the iteration, locking and degree control are checked on the CPU against LAPACK.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from scarf_amd import eigen


class _NumpyBlock:
    def __init__(self, M, n_top, b, seed=0):
        self.M, self.d, self.nb = M, M.shape[0], b
        self.ld = eigen._block_ld(b)
        self.Mw = M.copy()
        self.X = np.zeros((self.d, self.ld))
        self.X[:, :b] = np.random.default_rng(seed).standard_normal((self.d, b))
        self.L = np.zeros((0, self.d))
        self.theta_l = np.zeros(0)
        self.filter_degrees = []
        self.orthonormalize()

    def _cholqr_pass(self):
        n = self.nb
        X = self.X[:, :n]
        G = X.T @ X
        R = np.linalg.cholesky(0.5 * (G + G.T)).T
        self.X[:, :n] = X @ sla.solve_triangular(R, np.eye(n), lower=False)

    def orthonormalize(self):
        self._cholqr_pass()
        self._cholqr_pass()

    def orth_rayleigh_ritz(self):
        self._cholqr_pass()
        return self.rayleigh_ritz(generalized=True)

    def rayleigh_ritz(self, generalized=False):
        n = self.nb
        X = self.X[:, :n]
        XM = self.Mw @ X
        H = X.T @ XM
        G = X.T @ X if generalized else None
        w, Q = sla.eigh(0.5 * (H + H.T), None if G is None else 0.5 * (G + G.T))
        w, Q = w[::-1].copy(), Q[:, ::-1]
        self.X[:, :n] = X @ Q
        XM = XM @ Q
        res = np.linalg.norm(XM - self.X[:, :n] * w, axis=0)
        return w, res

    def lock(self, c, theta):
        Lnew = self.X[:, :c].T.copy()
        self.Mw = self.Mw - Lnew.T @ (np.asarray(theta)[:, None] * Lnew)
        self.L = np.vstack([self.L, Lnew])
        self.theta_l = np.concatenate([self.theta_l, theta])
        X = np.zeros_like(self.X)
        X[:, :self.nb - c] = self.X[:, c:self.nb]
        self.X = X
        self.nb -= c

    def filter(self, m, center, half, scale):
        self.filter_degrees.append(m)
        n = self.nb
        sig = half / (scale - center)
        tau = 2.0 / sig
        Y0 = self.X[:, :n]
        Y1 = (sig / half) * (self.Mw @ Y0 - center * Y0)
        for _ in range(2, m + 1):
            sn = 1.0 / (tau - sig)
            Y2 = (2 * sn / half) * (self.Mw @ Y1 - center * Y1) - sig * sn * Y0
            Y0, Y1, sig = Y1, Y2, sn
        self.X[:, :n] = Y1

    def result(self, n_active):
        Vt = np.vstack([self.L, self.X[:, :n_active].T])
        Y = Vt @ self.M
        th = np.einsum("ij,ij->i", Y, Vt)
        o = np.argsort(-th, kind="stable")
        th, Vt, Y = th[o], Vt[o], Y[o]
        res = float(np.max(np.linalg.norm(Y - th[:, None] * Vt, axis=1))) / th[0]
        orth = float(np.max(np.abs(Vt @ Vt.T - np.eye(Vt.shape[0]))))
        return th, Vt, res, orth


def _pca_like(d, n_signal, top_ratio, seed):
    """A merge-matrix-like spectrum: one dominant direction, a few signal components, and
    a Marchenko-Pastur-like bulk whose top edge holds the slow wanted pairs."""
    rng = np.random.default_rng(seed)
    bulk = np.sort(1.0 + 0.09 * np.cos(np.linspace(0, np.pi, d - n_signal - 1)) ** 3)[::-1]
    bulk = bulk * (1 + 1e-4 * rng.standard_normal(bulk.shape))
    sig = np.linspace(2.3, 1.5, n_signal)
    lam = np.concatenate([[top_ratio], sig, bulk]) * 1e6
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    M = (Q * lam) @ Q.T
    return 0.5 * (M + M.T)


@pytest.mark.parametrize("d,n_top,n_signal,top", [(500, 51, 19, 16.0), (700, 21, 5, 40.0),
                                                   (600, 101, 30, 8.0)])
def test_chebyshev_matches_lapack(d, n_top, n_signal, top):
    M = _pca_like(d, n_signal, top, seed=d)
    b = eigen.chebyshev_block_size(n_top, d)
    dev = _NumpyBlock(M, n_top, b)
    w, Vt, info = eigen.chebyshev_subspace(dev, n_top, 1e-12)
    wl, Vl = np.linalg.eigh(M)
    wl, Vl = wl[::-1][:n_top], Vl[:, ::-1][:, :n_top]
    assert info.max_residual <= 1e-10 and info.orthogonality <= 1e-10
    np.testing.assert_allclose(w, wl, rtol=0, atol=1e-11 * wl[0])
    # the dominant and signal components were deflated early, the bulk edge converged last
    assert dev.theta_l.shape[0] >= 1
    # eigenvectors up to sign, to the residual over the gap
    gaps = np.minimum(np.abs(np.diff(np.concatenate([[np.inf], wl]))),
                      np.abs(np.diff(np.concatenate([wl, [np.linalg.eigvalsh(M)[::-1][n_top]]]))))
    ov = np.abs(np.einsum("ij,ji->i", Vt, Vl))
    bound = 1 - 10 * (1e-12 * wl[0] / gaps) ** 2 - 1e-12
    assert np.all(ov >= np.minimum(bound, 1 - 1e-9))


@pytest.mark.parametrize("d,n,nt", [(600, 3000, 1), (800, 5000, 21)])
def test_chebyshev_gram_from_random_start(d, n, nt):
    """Plain Gram matrices (no dominant spike): the first Ritz values from the random
    block underestimate the spectrum's top, which the degree control must survive."""
    rng = np.random.default_rng(d)
    X = rng.standard_normal((n, d)) * np.geomspace(3, 1, d)
    M = X.T @ X
    dev = _NumpyBlock(M, nt, eigen.chebyshev_block_size(nt, d))
    w, Vt, info = eigen.chebyshev_subspace(dev, nt, 1e-12)
    wl = np.linalg.eigvalsh(M)[::-1][:nt]
    np.testing.assert_allclose(w, wl, rtol=0, atol=1e-11 * wl[0])
    assert info.max_residual <= 1e-10


def test_chebyshev_degree_keeps_block_conditioned():
    """The filtered block's condition number stays near e^dyn: the degree shrinks while
    the dominant direction is active and grows once it is locked."""
    M = _pca_like(400, 10, 60.0, seed=4)
    dev = _NumpyBlock(M, 31, eigen.chebyshev_block_size(31, 400))
    eigen.chebyshev_subspace(dev, 31, 1e-12)
    degs = dev.filter_degrees
    assert degs[0] <= 4 and max(degs) >= 20


def test_chebyshev_reports_nonconvergence():
    M = _pca_like(300, 5, 10.0, seed=5)
    dev = _NumpyBlock(M, 21, eigen.chebyshev_block_size(21, 300))
    with pytest.raises(eigen.EigenNotConverged):
        eigen.chebyshev_subspace(dev, 21, 1e-12, max_iter=2)


def test_block_layout():
    assert eigen.chebyshev_block_size(51, 2000) == 94
    assert eigen.chebyshev_block_size(101, 3000) == 187
    assert eigen.chebyshev_block_size(11, 40) == 27
    assert [eigen._block_ld(b) for b in (1, 16, 17, 94, 128, 129, 187, 256)] == \
        [16, 16, 32, 96, 128, 160, 192, 256]
