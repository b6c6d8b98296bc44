"""Host side of the GPU eigensolver: the block-tridiagonal bookkeeping.

The GPU kernels (sh_lanczos_steps) produce, per block step j, the diagonal block
A_j = V_j M V_j^T and the CholeskyQR factor R_j of the reorthogonalised next block.
A numpy restatement of those steps (same conventions: basis vectors as rows, W^T = Q R
with R upper triangular) feeds the host code here, which must give the Ritz pairs of
V M V^T and a residual bound that matches the true residual.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from scarf_amd import eigen

P = eigen.BLOCK


def _lanczos_numpy(M, J, seed=0):
    d = M.shape[0]
    rng = np.random.default_rng(seed)
    Q0, _ = np.linalg.qr(rng.standard_normal((d, P)))
    V = [Q0.T]
    T_diag, T_sub = [], []
    for j in range(J):
        W = V[j] @ M  # [P, d]: rows are M v (M symmetric)
        B = np.vstack(V)
        H = B @ W.T
        W = W - H.T @ B
        H2 = B @ W.T
        W = W - H2.T @ B
        T_diag.append((H + H2)[j * P:(j + 1) * P])
        Q, R = np.linalg.qr(W.T)
        s = np.sign(np.diag(R))  # CholeskyQR gives a positive diagonal
        Q, R = Q * s, R * s[:, None]
        T_sub.append(R)
        V.append(Q.T)
    return np.vstack(V), np.array(T_diag), np.array(T_sub)


def _spectrum_matrix(d, seed=1):
    rng = np.random.default_rng(seed)
    lam = np.concatenate([np.geomspace(50, 3, 12), 1 + 0.05 * rng.standard_normal(d - 12)])
    Qm, _ = np.linalg.qr(rng.standard_normal((d, d)))
    M = (Qm * lam) @ Qm.T
    return 0.5 * (M + M.T)


def test_lanczos_matrix_is_projection():
    d, J = 200, 10
    M = _spectrum_matrix(d)
    V, td, ts = _lanczos_numpy(M, J)
    k = J * P
    T = eigen.lanczos_matrix(td, ts, J)
    np.testing.assert_allclose(T, V[:k] @ M @ V[:k].T, atol=1e-10 * np.abs(M).max())
    # band storage round trip through LAPACK's banded solver
    ab = eigen._banded_upper(td, ts, J)
    w_band = sla.eig_banded(ab, lower=False, eigvals_only=True)
    np.testing.assert_allclose(w_band, np.linalg.eigvalsh(T), rtol=1e-12, atol=1e-12)


def test_residual_bound_matches_true_residual():
    d, J, nt = 300, 16, 12
    M = _spectrum_matrix(d, seed=3)
    V, td, ts = _lanczos_numpy(M, J)
    k = J * P
    ab = eigen._banded_upper(td, ts, J)
    th, S = sla.eig_banded(ab, lower=False, select="i", select_range=(k - nt, k - 1))
    th, S = th[::-1], S[:, ::-1]
    bound = np.linalg.norm(ts[J - 1] @ S[k - P:, :], axis=0)
    X = S.T @ V[:k]  # Ritz vectors as rows
    true = np.linalg.norm(X @ M - th[:, None] * X, axis=1)
    np.testing.assert_allclose(bound, true, rtol=1e-6, atol=1e-12 * th[0])
    w = np.linalg.eigvalsh(M)[::-1][:nt]
    # well separated top eigenvalues converge first
    np.testing.assert_allclose(th[:6], w[:6], rtol=1e-10)


@pytest.mark.parametrize("u", [1, 3, 8])
def test_band_storage(u):
    rng = np.random.default_rng(u)
    n = 30
    A = rng.standard_normal((n, n))
    A = A + A.T
    T = np.triu(np.tril(A, u), -u)
    ab = eigen._band_of(T, u)
    np.testing.assert_allclose(sla.eig_banded(ab, lower=False, eigvals_only=True),
                               np.linalg.eigvalsh(T), rtol=1e-12, atol=1e-12)


class _NumpyLanczos:
    """numpy stand-in for the GPU backend of thick_restart_lanczos (same conventions)."""

    def __init__(self, M, kin, seed=0):
        self.M, self.d = M, M.shape[0]
        rng = np.random.default_rng(seed)
        self.V = np.zeros((kin, self.d))
        self.V[:P] = np.linalg.qr(rng.standard_normal((self.d, P)))[0].T
        self.T_diag = np.zeros((kin // P, P, P))
        self.T_sub = np.zeros((kin // P, P, P))

    def steps(self, j0, j1):
        for j in range(j0, j1):
            k = (j + 1) * P
            W = self.V[j * P:k] @ self.M
            B = self.V[:k]
            H = B @ W.T
            W = W - H.T @ B
            H2 = B @ W.T
            W = W - H2.T @ B
            self.T_diag[j] = (H + H2)[j * P:k]
            Q, R = np.linalg.qr(W.T)
            sg = np.sign(np.diag(R))
            self.V[k:k + P] = (Q * sg).T
            self.T_sub[j] = R * sg[:, None]

    def read(self, j0, j1):
        return self.T_diag[j0:j1].copy(), self.T_sub[j0:j1].copy(), False

    def restart(self, S, k, m):
        Y = S.T @ self.V[:k]
        self.V[m:m + P] = self.V[k:k + P]
        self.V[:m] = Y

    def ritz(self, S, k):
        X = S.T @ self.V[:k]
        Y = X @ self.M
        th = np.einsum("ij,ij->i", Y, X)
        o = np.argsort(-th, kind="stable")
        th, X, Y = th[o], X[o], Y[o]
        res = np.linalg.norm(Y - th[:, None] * X, axis=1).max() / th[0]
        orth = np.abs(X @ X.T - np.eye(X.shape[0])).max()
        return th, X, res, orth


@pytest.mark.parametrize("d,nt", [(400, 10), (600, 31), (300, 1)])
def test_thick_restart_lanczos_host_loop(d, nt):
    """The restart loop (projected matrix diag(theta) + coupling b + new Lanczos blocks,
    residual bound ||R s_last||) converges to LAPACK's top eigenpairs."""
    M = _spectrum_matrix(d, seed=5)
    keep, kin = eigen.restart_sizes(nt, d)
    w, Vt, info = eigen.thick_restart_lanczos(_NumpyLanczos(M, kin), nt, keep, kin, 1e-12)
    w_ref, V_ref = sla.eigh(M, subset_by_index=[d - nt, d - 1])
    w_ref, V_ref = w_ref[::-1], V_ref[:, ::-1].T
    np.testing.assert_allclose(w, w_ref, rtol=1e-12)
    s = np.sign(np.sum(Vt * V_ref, axis=1))
    np.testing.assert_allclose(s[:, None] * Vt, V_ref, atol=1e-8)
    assert info.max_residual <= 1e-10 and info.cycles >= 1


def test_restart_sizes():
    keep, kin = eigen.restart_sizes(51, 2000)
    assert keep % P == 0 and kin % P == 0 and keep >= 51 + 16 and kin >= 2 * keep
    with pytest.raises(eigen.EigenNotConverged):
        eigen.restart_sizes(51, 100)
