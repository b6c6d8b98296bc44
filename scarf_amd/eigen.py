"""Top eigenpairs of the symmetric FP64 PCA merge matrix, on the GPU.

``IncrementalPCA.partial_fit`` (scikit-learn ``_incremental_pca.py:345-363``, called
from scarf ``ann.py:241,250``) takes the SVD of the stacked rows
``[V^T S; X - m_c; mean correction]``; its right singular vectors and squared singular
values are the eigenpairs of the d x d merge matrix
``M = V^T S^2 V + C + m_c m_c^T`` that ``sh_merge_matrix`` forms on the GPU.

The wanted ``n_top = dims + 1`` eigenpairs come from block Lanczos (block size 8,
``sh_lanczos_steps``) with full reorthogonalisation: every d-sized operation runs on
the GPU; the basis is restarted every cycle (thick restart: the best Ritz vectors and
the newest block are kept), so the host only solves projected matrices of about
4 n_top rows.  It stops when every wanted pair has a residual
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

BLOCK = 8


@dataclasses.dataclass
class EighInfo:
    vectors: int  # basis size of the last cycle
    cycles: int
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

    Thick-restart block Lanczos: the GPU grows the basis by block steps; every cycle the
    host solves the small projected matrix, and the GPU restarts the basis as
    [best Ritz vectors; newest block] (one FP64 GEMM).  Returns ``(w, Vt, info)``:
    eigenvalues descending (host f64 [n_top]), eigenvectors as rows (host f64
    [n_top, d]) and an :class:`EighInfo`.  Raises :class:`EigenNotConverged` when the
    basis is too small for the request or the cycles run out.
    """
    require_gpu()
    d = int(M.shape[0])
    if M.dtype != torch.float64 or M.shape != (d, d) or not M.is_contiguous():
        raise ValueError("M must be a contiguous float64 [d, d] tensor")
    if not 1 <= n_top <= d:
        raise ValueError(f"n_top={n_top} must be in [1, {d}]")
    keep, kin = restart_sizes(n_top, d, max_vectors)
    dev = _DeviceLanczos(M, kin, seed, stream)
    return thick_restart_lanczos(dev, n_top, keep, kin, tol)


def restart_sizes(n_top: int, d: int, max_vectors: Optional[int] = None):
    """(Ritz vectors kept per restart, basis size per cycle), multiples of 8."""
    P = BLOCK
    keep = int(math.ceil(max(n_top + 16, 1.8 * n_top) / P)) * P
    kin = max(2 * keep, keep + 8 * P) + P
    cap = (d // P) * P
    if max_vectors is not None:
        cap = min(cap, (int(max_vectors) // P) * P)
    kin = min(kin, cap)
    if kin < keep + 3 * P:
        raise EigenNotConverged(f"basis capacity {cap} too small for {n_top} pairs of d={d}")
    return keep, kin


def thick_restart_lanczos(dev, n_top: int, keep: int, kin: int, tol: float,
                          max_cycles: int = 200):
    """Host loop of thick-restart block Lanczos over a device backend ``dev`` providing
    ``steps(j0, j1)``, ``read(j0, j1) -> (T_diag[j0:j1], T_sub[j0:j1], broke)``,
    ``restart(S, k, m)`` and ``ritz(S, k) -> (w, Vt, res, orth)``.

    After a restart the basis is [Y (m Ritz vectors); q (the newest block, block m/8)]
    with the projected matrix diag(theta) coupled to q through b = R_{J-1} S_last; the
    residual of a Ritz pair (theta, s) of the projected matrix is ||R_{J-1} s_last||.
    """
    P = BLOCK
    theta = np.zeros(0)
    b = np.zeros((P, 0))
    J0 = 0  # block index of q (the first block whose step runs in this cycle)
    res_bound, scale = np.array([np.inf]), 1.0
    for cycle in range(1, max_cycles + 1):
        J = kin // P - 1  # step j writes block j + 1 < kin / 8
        dev.steps(J0, J)
        td, ts, broke = dev.read(J0, J)
        if broke:
            raise EigenNotConverged(f"block Lanczos broke down in cycle {cycle}")
        m = theta.shape[0]
        k = J * P
        H = np.zeros((k, k))
        if m:
            H[:m, :m] = np.diag(theta)
            H[m:m + P, :m] = b
            H[:m, m:m + P] = b.T
        H[m:, m:] = lanczos_matrix(td, ts, J - J0)
        w_all, Z = np.linalg.eigh(H)
        mk = min(keep, k)
        theta = w_all[::-1][:mk].copy()
        S = np.ascontiguousarray(Z[:, ::-1][:, :mk])
        R_last = ts[-1]
        res_bound = np.linalg.norm(R_last @ S[k - P:k, :n_top], axis=0)
        scale = max(abs(theta[0]), np.finfo(float).tiny)
        if float(res_bound.max()) <= tol * scale:
            w, Vt, res, orth = dev.ritz(S[:, :n_top], k)
            if res <= 100.0 * tol and orth <= 1e-10:
                return w, Vt, EighInfo(k, cycle, res, orth)
        b = R_last @ S[k - P:k, :]
        dev.restart(S, k, mk)
        J0 = mk // P
    raise EigenNotConverged(f"no convergence in {max_cycles} restart cycles "
                            f"(residual {float(res_bound.max()) / scale:.2e})")


class _DeviceLanczos:
    """GPU backend of thick_restart_lanczos: basis, Lanczos blocks and restarts on the
    device through sh_lanczos_steps and sh_dgemm."""

    def __init__(self, M, kin: int, seed: int, stream):
        P = BLOCK
        self.M, self.d, self.kin = M, int(M.shape[0]), kin
        dv = M.device
        self.st = stream_handle(stream)
        self.V = torch.empty((kin, self.d), dtype=torch.float64, device=dv)
        rng = np.random.default_rng(seed)
        Q0, _ = np.linalg.qr(rng.standard_normal((self.d, P)))
        self.V[:P].copy_(torch.from_numpy(np.ascontiguousarray(Q0.T)))
        nblk = kin // P
        self.T_diag = torch.zeros((nblk, P, P), dtype=torch.float64, device=dv)
        self.T_sub = torch.zeros((nblk, P, P), dtype=torch.float64, device=dv)
        self.status = torch.zeros(1, dtype=torch.int32, device=dv)
        self.Y = torch.empty((kin, self.d), dtype=torch.float64, device=dv)
        self.buf, self.nbytes = workspace("lanczos").get(
            _lib.lib().sh_lanczos_workspace(self.d, kin))

    def steps(self, j0, j1):
        if j1 > j0:
            _lib.call("sh_lanczos_steps", ptr(self.M), self.d, ptr(self.V), self.kin, j0,
                      j1 - j0, ptr(self.T_diag), ptr(self.T_sub), ptr(self.status),
                      ptr(self.buf), self.nbytes, self.st)

    def read(self, j0, j1):
        td = self.T_diag[j0:j1].cpu().numpy()
        ts = self.T_sub[j0:j1].cpu().numpy()
        return td, ts, int(self.status.item()) != 0

    def restart(self, S, k, m):
        P = BLOCK
        S_dev = torch.from_numpy(np.array(S, order="C", copy=True)).to(self.V.device)
        _lib.call("sh_dgemm", 1, m, self.d, k, 1.0, ptr(S_dev), m, ptr(self.V), self.d, 0.0,
                  ptr(self.Y), self.d, self.st)
        self.V[m:m + P].copy_(self.V[k:k + P])  # k >= m + 8: the row ranges are disjoint
        self.V[:m].copy_(self.Y[:m])

    def ritz(self, S, k):
        return _ritz_and_verify(self.M, self.V, S, k, self.d, S.shape[1], self.st)


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
