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
