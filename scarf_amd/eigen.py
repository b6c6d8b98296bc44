"""Top eigenpairs of the symmetric FP64 PCA merge matrix, on the GPU.

``IncrementalPCA.partial_fit`` (scikit-learn ``_incremental_pca.py:345-363``, called
from scarf ``ann.py:241,250``) takes the SVD of the stacked rows
``[V^T S; X - m_c; mean correction]``; its right singular vectors and squared singular
values are the eigenpairs of the d x d merge matrix
``M = V^T S^2 V + C + m_c m_c^T`` that ``sh_merge_matrix`` forms on the GPU.

The wanted ``n_top = dims + 1`` eigenpairs come from block Lanczos (block size 8,
``sh_lanczos_steps``) with full reorthogonalisation: every d-sized operation runs on
the GPU; the host solves the small banded Lanczos matrix (at most a few thousand
rows) for its Ritz pairs and stops when every wanted pair has a residual
``||M x - theta x|| <= tol * theta_max`` (a bound read off the last block, then
verified on the Ritz vectors themselves with one GEMM).  The eigenvectors then agree
with LAPACK's to within ``tol / relative gap``; eigenvalues are the Rayleigh
quotients of the returned vectors.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

from . import _lib
from .device import ptr, require_gpu, stream_handle, workspace

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

try:
    import scipy.linalg as _sla
except ImportError:  # pragma: no cover
    _sla = None

BLOCK = 8


@dataclasses.dataclass
class EighInfo:
    vectors: int  # Krylov basis size at convergence
    checks: int
    max_residual: float  # max ||M x - theta x|| / theta_max over the wanted pairs
    orthogonality: float  # max |Vt Vt^T - I|


class EigenNotConverged(RuntimeError):
    """The basis reached its capacity before the wanted Ritz pairs converged."""


def _banded_upper(T_diag: np.ndarray, T_sub: np.ndarray, J: int) -> np.ndarray:
    """Upper band storage (bandwidth 8) of the block-tridiagonal Lanczos matrix of J blocks."""
    return _band_of(lanczos_matrix(T_diag, T_sub, J), BLOCK)


def lanczos_matrix(T_diag: np.ndarray, T_sub: np.ndarray, J: int) -> np.ndarray:
    """Dense block-tridiagonal T (J*8 square): diagonal blocks A_j (symmetrised), R_j
    below the diagonal and R_j^T above it."""
    P = BLOCK
    T = np.zeros((J * P, J * P))
    for j in range(J):
        A = T_diag[j]
        T[j * P:(j + 1) * P, j * P:(j + 1) * P] = 0.5 * (A + A.T)
        if j + 1 < J:
            T[(j + 1) * P:(j + 2) * P, j * P:(j + 1) * P] = T_sub[j]
            T[j * P:(j + 1) * P, (j + 1) * P:(j + 2) * P] = T_sub[j].T
    return T


def _band_of(T: np.ndarray, u: int) -> np.ndarray:
    """LAPACK upper band storage: ab[u + i - j, j] = T[i, j] for j - u <= i <= j."""
    k = T.shape[0]
    ab = np.zeros((u + 1, k))
    for r in range(u + 1):
        ab[u - r, r:] = np.diagonal(T, r)
    return ab


def top_eigh_device(M: "torch.Tensor", n_top: int, tol: float = 1e-12,
                    max_vectors: Optional[int] = None, seed: int = 0, stream=None):
    """Top ``n_top`` eigenpairs of the symmetric FP64 device matrix ``M`` (d x d).

    Returns ``(w, Vt, info)``: eigenvalues descending (host f64 [n_top]), eigenvectors as
    rows (host f64 [n_top, d]) and an :class:`EighInfo`.  Raises
    :class:`EigenNotConverged` when the basis capacity is reached first.
    """
    require_gpu()
    if _sla is None:
        raise RuntimeError("scipy is required for the banded Ritz solve")
    d = int(M.shape[0])
    if M.dtype != torch.float64 or M.shape != (d, d) or not M.is_contiguous():
        raise ValueError("M must be a contiguous float64 [d, d] tensor")
    if not 1 <= n_top <= d:
        raise ValueError(f"n_top={n_top} must be in [1, {d}]")
    P = BLOCK
    cap = (d // P) * P
    kmax = cap if max_vectors is None else min(cap, (int(max_vectors) // P) * P)
    if kmax < max(2 * P, n_top + P):
        raise EigenNotConverged(f"basis capacity {kmax} too small for {n_top} pairs of d={d}")
    dev = M.device
    st = stream_handle(stream)
    L = _lib.lib()
    V = torch.empty((kmax, d), dtype=torch.float64, device=dev)
    rng = np.random.default_rng(seed)
    Q0, _ = np.linalg.qr(rng.standard_normal((d, P)))
    V[:P].copy_(torch.from_numpy(np.ascontiguousarray(Q0.T)))
    nblk = kmax // P
    T_diag = torch.zeros((nblk, P, P), dtype=torch.float64, device=dev)
    T_sub = torch.zeros((nblk, P, P), dtype=torch.float64, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    buf, nbytes = workspace("lanczos").get(L.sh_lanczos_workspace(d, kmax))

    J = 0  # blocks whose step has run
    target = max(4 * n_top, 8 * P)
    checks = 0
    while True:
        # steps until the basis reaches the next check point (or the capacity)
        J_next = min(nblk - 1, max(J + 1, int(math.ceil(target / P))))
        _lib.call("sh_lanczos_steps", ptr(M), d, ptr(V), kmax, J, J_next - J, ptr(T_diag),
                  ptr(T_sub), ptr(status), ptr(buf), nbytes, st)
        J = J_next
        checks += 1
        td = T_diag[:J].cpu().numpy()
        ts = T_sub[:J].cpu().numpy()
        broke = int(status.item()) != 0
        k = J * P
        ab = _banded_upper(td, ts, J)
        nt = min(n_top, k)
        th, S = _sla.eig_banded(ab, lower=False, select="i", select_range=(k - nt, k - 1),
                                check_finite=False)
        th, S = th[::-1], S[:, ::-1]
        res_bound = np.linalg.norm(ts[J - 1] @ S[k - P:, :], axis=0)
        scale = max(abs(th[0]), np.finfo(float).tiny)
        converged = nt == n_top and not broke and float(res_bound.max()) <= tol * scale
        if converged:
            w, Vt, res, orth = _ritz_and_verify(M, V, S, k, d, n_top, st)
            if res <= 100.0 * tol and orth <= 1e-10:
                return w, Vt, EighInfo(k, checks, res, orth)
        if broke or J >= nblk - 1:
            raise EigenNotConverged(
                f"block Lanczos stopped at {k} of {kmax} vectors (breakdown={broke}, "
                f"residual {float(res_bound.max()) / scale:.2e})")
        target = int(k * 1.25) + P


def _ritz_and_verify(M, V, S, k, d, n_top, st):
    dev = M.device
    S_dev = torch.from_numpy(np.array(S, order="C", copy=True)).to(dev)  # [k, n_top]
    Vt = torch.empty((n_top, d), dtype=torch.float64, device=dev)
    _lib.call("sh_dgemm", 1, n_top, d, k, 1.0, ptr(S_dev), n_top, ptr(V), d, 0.0, ptr(Vt), d,
              st)
    Y = torch.empty((n_top, d), dtype=torch.float64, device=dev)
    _lib.call("sh_dgemm", 0, n_top, d, d, 1.0, ptr(Vt), d, ptr(M), d, 0.0, ptr(Y), d, st)
    Vh = Vt.cpu().numpy()
    Yh = Y.cpu().numpy()
    theta = np.einsum("ij,ij->i", Yh, Vh)
    order = np.argsort(-theta, kind="stable")
    theta, Vh, Yh = theta[order], Vh[order], Yh[order]
    scale = max(abs(theta[0]), np.finfo(float).tiny)
    res = float(np.max(np.linalg.norm(Yh - theta[:, None] * Vh, axis=1))) / scale
    orth = float(np.max(np.abs(Vh @ Vh.T - np.eye(n_top))))
    return theta, Vh, res, orth
