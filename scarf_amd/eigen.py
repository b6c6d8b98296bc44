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

import contextlib

import numpy as np
import scipy.linalg as _sla

from . import _lib
from .device import ptr, require_gpu, stream_handle, workspace

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

try:
    from threadpoolctl import ThreadpoolController as _ThreadpoolController
except ImportError:  # pragma: no cover
    _ThreadpoolController = None
_controller = None


def _one_thread():
    """The host's small (b x b, b <= 256) LAPACK problems run single-threaded: on a 16-thread
    host a 94 x 94 symmetric eigensolve takes 0.36 ms on one thread (divide and conquer)
    against 0.6 ms with the BLAS library's default thread count.  One controller is built
    once (it scans the loaded libraries); its limit() then costs microseconds."""
    global _controller
    if _ThreadpoolController is None:
        return contextlib.nullcontext()
    if _controller is None:
        _controller = _ThreadpoolController()
    return _controller.limit(limits=1)

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
    with torch.cuda.stream(stream):  # read-backs on the kernels' stream
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
        with _one_thread():
            w_all, Z = _sla.eigh(H, driver="evd")
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


# ---- block subspace iteration with Chebyshev filtering ---------------------------------

def chebyshev_block_size(n_top: int, d: int) -> int:
    """Active block of the filtered subspace iteration: the wanted pairs plus about 0.85
    as many guard vectors (at least 16).  Convergence of the last wanted pair is set by
    its distance to the block's smallest Ritz value, not to its neighbour."""
    b = n_top + max(16, int(0.85 * n_top + 0.5))
    return min(b, d)


def _block_ld(b: int) -> int:
    """Columns of the device block arrays: 16-aligned, and a multiple of 32 past 128 so
    that sh_symm_block splits them into equal groups of at most 128."""
    return -(-b // 16) * 16 if b <= 128 else -(-b // 32) * 32


def chebyshev_subspace(dev, n_top: int, tol: float, dyn: float = 16.0,
                       lock_factor: float = 0.1, max_degree: int = 80,
                       max_iter: int = 100):
    """Host loop of Chebyshev-filtered block subspace iteration over a backend ``dev``
    providing ``rayleigh_ritz() -> (theta, res)`` (active block, descending),
    ``lock(c, theta)``, ``filter(m, center, half_width, scale)``,
    ``orth_rayleigh_ritz() -> (theta, res)`` (one CholeskyQR pass, then the Ritz pairs of
    the block from H q = theta G q) and ``result(n_active) -> (w, Vt, res, orth)``.

    Each iteration filters the active block with a Chebyshev polynomial that damps
    [0, cut] (the merge matrix is positive semi-definite; cut is the block's smallest
    Ritz value) and is scaled to 1 at an upper bound of the active spectrum, then takes
    one CholeskyQR pass and the Ritz pairs of the generalized problem H q = theta G q (G the
    block's Gram after that pass, so a second pass is not needed).  The degree is chosen so
    that the filtered block's condition number stays below e^dyn.  Leading pairs whose
    residual is far below the tolerance (so that removing them perturbs the rest by less
    than lock_factor * tol * theta_max) are locked and deflated out of the working
    matrix, which shrinks the dynamic range of the next filters: the well separated top
    components go first, the slow ones near the noise edge converge last.
    """
    theta_max = 0.0
    locked = 0
    lam_up = None
    pending = None  # Ritz pairs of the active block still valid after a lock
    for it in range(1, max_iter + 1):
        th, res = dev.rayleigh_ritz() if pending is None else pending
        pending = None
        need = n_top - locked
        theta_max = max(theta_max, float(th[0]))
        thr = tol * theta_max
        if np.all(res[:need] <= thr):
            w, Vt, res_v, orth = dev.result(need)
            if res_v <= 100.0 * tol and orth <= 1e-10:
                return w, Vt, EighInfo(int(th.shape[0]) + locked, it, res_v, orth)
        c = 0
        for cand in range(need, 0, -1):
            gap = float(th[cand - 1] - th[cand])
            if gap > 0 and np.all(res[:cand] * th[:cand] / gap <= lock_factor * thr):
                c = cand
                break
        if c == need:  # every wanted pair is final; the check above decides
            c = 0
        if c:
            # the rest of the block is already a Ritz basis orthogonal to the locked
            # vectors, so deflating them leaves its Ritz pairs as they are
            dev.lock(c, th[:c])
            locked += c
            lam_up = None
            pending = (th[c:], res[c:])
            continue
        # upper end of the active spectrum: the top Ritz value plus its residual, the
        # largest seen since the last lock (early Ritz values underestimate); until the
        # top pair has settled the degree stays low so an underestimate cannot blow up
        # the block's condition number
        up = float(th[0] + res[0])
        lam_up = up if lam_up is None else max(lam_up, up)
        cut = float(th[-1])
        if not (cut > 0.0 and lam_up > cut):
            raise EigenNotConverged(f"degenerate filter interval (cut {cut:.3e}, top {lam_up:.3e})")
        half = 0.5 * cut
        x_up = (lam_up - half) / half
        m = int(dyn / math.acosh(max(x_up, 1.0 + 1e-12)))
        m = max(2, min(max_degree if res[0] <= 1e-3 * th[0] else 4, m))
        dev.filter(m, half, half, lam_up)
        # one CholeskyQR pass, then the Rayleigh-Ritz step as H q = theta G q with the
        # block's Gram G (close to I after that pass): the second CholeskyQR pass folded
        # into the small eigenproblem, one host round trip fewer per iteration
        pending = dev.orth_rayleigh_ritz()
    raise EigenNotConverged(f"no convergence in {max_iter} filtered iterations")


def top_eigh_chebyshev(M: "torch.Tensor", n_top: int, tol: float = 1e-12, seed: int = 0,
                       stream=None, block: Optional[int] = None,
                       start: Optional[np.ndarray] = None):
    """Top ``n_top`` eigenpairs of the symmetric positive semi-definite FP64 device matrix
    ``M`` (d x d) by Chebyshev-filtered block subspace iteration (see
    :func:`chebyshev_subspace`).  Every d-sized operation runs on the GPU
    (``sh_symm_block``, ``sh_block_gram``, ``sh_dgemm``, ``sh_block_residual``); the host
    solves the b x b Rayleigh-Ritz and Cholesky problems.  ``start`` (host, k x d rows)
    seeds the first k block vectors, e.g. the previous partial fit's components; the rest
    of the block starts random.  Returns ``(w, Vt, info)`` like :func:`top_eigh_device`."""
    require_gpu()
    d = int(M.shape[0])
    if M.dtype != torch.float64 or M.shape != (d, d) or not M.is_contiguous():
        raise ValueError("M must be a contiguous float64 [d, d] tensor")
    if not 1 <= n_top <= d:
        raise ValueError(f"n_top={n_top} must be in [1, {d}]")
    b = chebyshev_block_size(n_top, d) if block is None else int(block)
    if b <= n_top or b > 256 or b > d:
        raise EigenNotConverged(f"block {b} cannot hold {n_top} pairs plus guards (d={d})")
    # torch's copies and read-backs on the stream the library kernels use (a no-op context
    # for the current stream)
    with torch.cuda.stream(stream):
        dev = _DeviceBlock(M, n_top, b, seed, stream, start)
        return chebyshev_subspace(dev, n_top, tol)


_RANDOM_BLOCKS = {}


def _random_block(d: int, b: int, seed: int) -> np.ndarray:
    """The seeded random start block (d x b), drawn once per shape: every partial fit of
    a run solves at the same size."""
    key = (d, b, seed)
    blk = _RANDOM_BLOCKS.get(key)
    if blk is None:
        blk = _RANDOM_BLOCKS[key] = np.random.default_rng(seed).standard_normal((d, b))
    return blk


class _DeviceBlock:
    """GPU backend of chebyshev_subspace.  The active block is the first ``nb`` columns
    of d x ld arrays (the rest zero); locked vectors are kept as rows and deflated out of
    a working copy of M (Mw -= L^T diag(theta) L, two FP64 GEMMs)."""

    def __init__(self, M, n_top: int, b: int, seed: int, stream, start=None):
        self.M, self.d, self.nb = M, int(M.shape[0]), b
        self.ld = _block_ld(b)
        dv = M.device
        self.st = stream_handle(stream)
        self.Mw = M.clone()
        self.buf = [torch.zeros((self.d, self.ld), dtype=torch.float64, device=dv)
                    for _ in range(4)]
        self.X = 0  # index of the block in self.buf
        self.L = torch.empty((n_top, self.d), dtype=torch.float64, device=dv)
        self.T = torch.empty((n_top, self.d), dtype=torch.float64, device=dv)
        self.theta_l = np.zeros(0)
        self.th_dev = torch.zeros(self.ld, dtype=torch.float64, device=dv)
        self.res_dev = torch.zeros(self.ld, dtype=torch.float64, device=dv)
        self.G = torch.empty((self.ld, self.ld), dtype=torch.float64, device=dv)
        L = _lib.lib()
        # the k-slice count grows as the block shrinks: size for the narrowest block
        self.ws_sb, self.nb_sb = workspace("chfsi_sb").get(
            max(L.sh_symm_block_workspace(self.d, c, self.ld) for c in range(16, self.ld + 1, 16)))
        self.ws_g, self.nb_g = workspace("chfsi_gram").get(
            L.sh_block_gram_workspace(self.d, self.ld, self.ld))
        X0 = _random_block(self.d, b, seed)
        if start is not None:
            X0 = X0.copy()
            k = min(int(start.shape[0]), b - 1)
            X0[:, :k] = np.asarray(start, dtype=np.float64)[:k].T
        self.buf[0][:, :b].copy_(torch.from_numpy(X0))
        self.orthonormalize()

    def _cols(self) -> int:
        return _block_ld(self.nb)

    def _gram(self, A, B, n):
        _lib.call("sh_block_gram", ptr(A), self.ld, n, ptr(B), self.ld, n, self.d, ptr(self.G),
                  ptr(self.ws_g), self.nb_g, self.st)
        return self.G.view(-1)[:n * n].view(n, n)

    def _times_small(self, A, S: np.ndarray, out):
        """out[:, :n] = A[:, :k] S for a host matrix S (k x n)."""
        k, n = S.shape
        S_dev = torch.from_numpy(np.ascontiguousarray(S)).to(self.M.device)
        _lib.call("sh_dgemm", 0, self.d, n, k, 1.0, ptr(A), self.ld, ptr(S_dev), n, 0.0,
                  ptr(out), self.ld, self.st)

    def _spare(self, *used):
        for i in range(len(self.buf)):
            if i not in used:
                return i
        raise RuntimeError("no spare block buffer")

    def _cholqr_pass(self):
        n = self.nb
        X = self.buf[self.X]
        G = self._gram(X, X, n).cpu().numpy()
        G = 0.5 * (G + G.T)
        try:
            with _one_thread():
                R = _sla.cholesky(G, lower=False)  # G = R^T R (scipy: see rayleigh_ritz)
        except np.linalg.LinAlgError as exc:
            raise EigenNotConverged("filtered block lost rank (CholeskyQR)") from exc
        with _one_thread():
            Rinv, info = _sla.lapack.dtrtri(R, lower=0)
        if info != 0:
            raise EigenNotConverged("filtered block lost rank (CholeskyQR)")
        o = self._spare(self.X)
        self._times_small(X, Rinv, self.buf[o])
        self.X = o

    def orthonormalize(self):
        self._cholqr_pass()
        self._cholqr_pass()

    def orth_rayleigh_ritz(self):
        self._cholqr_pass()
        return self.rayleigh_ritz(generalized=True)

    def rayleigh_ritz(self, generalized: bool = False):
        n, cols = self.nb, self._cols()
        X = self.buf[self.X]
        xm = self._spare(self.X)
        XM = self.buf[xm]
        _lib.call("sh_symm_block", ptr(self.Mw), self.d, ptr(X), None, cols, self.ld, 1.0, 0.0,
                  0.0, ptr(XM), ptr(self.ws_sb), self.nb_sb, self.st)
        Hd = self._gram(X, XM, n).clone()
        if generalized:  # the block's own Gram, read back with H in one copy
            Gd = self._gram(X, X, n)
            HG = torch.cat([Hd, Gd]).cpu().numpy()
            H, G = HG[:n], HG[n:]
        else:
            H, G = Hd.cpu().numpy(), None
        # scipy's LAPACK binding: numpy's eigh and cholesky are an order of magnitude
        # slower at these sizes on a many-threaded host
        try:
            with _one_thread():
                w, Q = _sla.eigh(0.5 * (H + H.T), None if G is None else 0.5 * (G + G.T),
                                 driver="evd" if G is None else "gvd")
        except np.linalg.LinAlgError as exc:
            raise EigenNotConverged("filtered block lost rank (Rayleigh-Ritz)") from exc
        w, Q = w[::-1].copy(), np.ascontiguousarray(Q[:, ::-1])
        xo = self._spare(self.X, xm)
        mo = self._spare(self.X, xm, xo)
        self._times_small(X, Q, self.buf[xo])
        self._times_small(XM, Q, self.buf[mo])
        self.th_dev[:n].copy_(torch.from_numpy(w))
        _lib.call("sh_block_residual", ptr(self.buf[xo]), ptr(self.buf[mo]), self.d, n, self.ld,
                  ptr(self.th_dev), ptr(self.res_dev), self.st)
        self.X = xo
        return w, self.res_dev[:n].cpu().numpy()

    def lock(self, c: int, theta: np.ndarray):
        nl = self.theta_l.shape[0]
        X = self.buf[self.X]
        L = self.L[nl:nl + c]
        L.copy_(X[:, :c].t())
        D = torch.from_numpy(np.diag(np.asarray(theta, dtype=np.float64))).to(self.M.device)
        T = self.T[:c]
        _lib.call("sh_dgemm", 0, c, self.d, c, 1.0, ptr(D), c, ptr(L), self.d, 0.0, ptr(T),
                  self.d, self.st)
        _lib.call("sh_dgemm", 1, self.d, self.d, c, -1.0, ptr(L), self.d, ptr(T), self.d, 1.0,
                  ptr(self.Mw), self.d, self.st)
        self.theta_l = np.concatenate([self.theta_l, np.asarray(theta, dtype=np.float64)])
        o = self._spare(self.X)
        Y = self.buf[o]
        Y.zero_()
        Y[:, :self.nb - c].copy_(X[:, c:self.nb])
        self.X = o
        self.nb -= c
        for i, t in enumerate(self.buf):  # columns past the active block stay zero
            if i != o:
                t[:, self.nb:].zero_()

    def filter(self, m: int, center: float, half: float, scale: float):
        cols = self._cols()
        sig = half / (scale - center)
        tau = 2.0 / sig
        y0 = self.X
        y1 = self._spare(y0)
        _lib.call("sh_symm_block", ptr(self.Mw), self.d, ptr(self.buf[y0]), None, cols, self.ld,
                  sig / half, -center * sig / half, 0.0, ptr(self.buf[y1]), ptr(self.ws_sb),
                  self.nb_sb, self.st)
        for _ in range(2, m + 1):
            sn = 1.0 / (tau - sig)
            y2 = self._spare(y0, y1)
            _lib.call("sh_symm_block", ptr(self.Mw), self.d, ptr(self.buf[y1]),
                      ptr(self.buf[y0]), cols, self.ld, 2.0 * sn / half, -2.0 * sn * center / half,
                      -sig * sn, ptr(self.buf[y2]), ptr(self.ws_sb), self.nb_sb, self.st)
            y0, y1, sig = y1, y2, sn
        self.X = y1

    def result(self, n_active: int):
        nl = self.theta_l.shape[0]
        n_top = nl + n_active
        Vt = torch.empty((n_top, self.d), dtype=torch.float64, device=self.M.device)
        if nl:
            Vt[:nl].copy_(self.L[:nl])
        Vt[nl:].copy_(self.buf[self.X][:, :n_active].t())
        return _verify_rows(self.M, Vt, self.d, self.st)


def _verify_rows(M, Vt, d, st):
    """Rayleigh quotients of the rows of Vt (device, n x d) under M, sorted descending,
    with the largest relative residual and the orthogonality defect (one FP64 GEMM)."""
    n_top = int(Vt.shape[0])
    Y = torch.empty((n_top, d), dtype=torch.float64, device=M.device)
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
