"""PCA fit (incremental, in scarf's block order) and the sparse projection.

Fit: for each ``partial_fit`` call of scarf's ``AnnStream._fit_pca``
(``ann.py:207-256``: held-back first block fitted last, small blocks carried over,
``LinAlgError`` blocks carried over) the GPU densifies the call's rows as
z = (x - mu) / sigma and accumulates Z^T Z and the column sums on MFMA
(``sh_densify`` + ``sh_gram``).  The host then restates
``IncrementalPCA.partial_fit`` (scikit-learn ``_incremental_pca.py:257-385``) in FP64:
the chunk-centred Gram ``C = Z^T Z - n m m^T``, the Chan mean/variance update, the
merge ``V^T S^2 V + C + m_c m_c^T`` and the top-(dims+1) symmetric eigensolve
(LAPACK), with scikit-learn's ``svd_flip(u_based_decision=False)`` sign rule.

``mode="exact"`` fits one Gram over all selected rows (exact PCA of z); it equals
the incremental fit when there is a single block, and needs one eigensolve.

Projection (``AnnStream.reducer``, ``ann.py:129-138``): E = z W evaluated sparsely
on the GPU (``sh_project``).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

from . import _lib
from .device import DeviceCSR, ptr, require_gpu, stream_handle, workspace
from .eigen import EigenNotConverged, top_eigh_chebyshev, top_eigh_device
from .normalize import FeatureStats, NormParams, chunk_bounds

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

try:
    import scipy.linalg as _sla
except ImportError:  # pragma: no cover
    _sla = None


# below this many features the host LAPACK solve takes a few milliseconds and the
# Krylov basis would span most of the space
EIG_DEVICE_MIN_D = 512

# the sparse Gram forms sum_r nnz_r^2 / 2 products with LDS atomics where the dense one
# forms n d^2 / 2 on FP64 MFMA; it is taken up to this density (nnz / (n d))
SPARSE_GRAM_MAX_DENSITY = 0.3
SPARSE_GRAM_MAX_D = 8063


def _top_eigh(M: np.ndarray, n_top: int):
    """Top ``n_top`` eigenpairs of a symmetric FP64 matrix, descending."""
    d = M.shape[0]
    n_top = min(n_top, d)
    if _sla is not None and n_top < d:
        w, V = _sla.eigh(M, subset_by_index=[d - n_top, d - 1], driver="evr",
                         check_finite=False)
    else:
        w, V = np.linalg.eigh(M)
        w, V = w[d - n_top:], V[:, d - n_top:]
    return w[::-1].copy(), V[:, ::-1].T.copy()


def _flip_rows(Vt: np.ndarray) -> np.ndarray:
    """scikit-learn svd_flip(u_based_decision=False): max-|.| entry of each row > 0."""
    idx = np.argmax(np.abs(Vt), axis=1)
    signs = np.sign(Vt[np.arange(Vt.shape[0]), idx])
    return Vt * signs[:, None]


class IPCAState:
    """The fitted attributes of ``IncrementalPCA`` that scarf and its callers read."""

    def __init__(self, n_components: int):
        self.n_components = int(n_components)
        self.n_components_ = self.n_components
        self.n_samples_seen_ = 0
        self.mean_ = 0.0
        self.var_ = 0.0
        self.components_: Optional[np.ndarray] = None
        self.singular_values_: Optional[np.ndarray] = None
        self.explained_variance_: Optional[np.ndarray] = None
        self.explained_variance_ratio_: Optional[np.ndarray] = None
        self.noise_variance_ = 0.0
        self.eig_log: list = []  # (backend, basis size or d) per partial fit

    def merge(self, n: int, G, colsum: np.ndarray, eig: str = "auto", stream=None):
        """One partial_fit from the uncentred Gram and column sums of the chunk's z.

        ``G`` is a host FP64 array or a device FP64 tensor (then the merge matrix is
        formed on the GPU).  ``eig``: "device" (GPU Chebyshev-filtered block subspace
        iteration), "lanczos" (GPU thick-restart block Lanczos), "host" (LAPACK) or "auto"
        (device for a device ``G`` with d >= EIG_DEVICE_MIN_D, LAPACK when the device solve
        does not converge).
        """
        on_dev = torch is not None and isinstance(G, torch.Tensor)
        d = int(G.shape[0])
        if self.components_ is None and self.n_components > n:
            raise ValueError(f"n_components={self.n_components} must be <= the first batch's "
                             f"sample count {n}")
        if self.n_components > d:
            raise ValueError(f"n_components={self.n_components} invalid for n_features={d}")
        if eig not in ("auto", "device", "lanczos", "host"):
            raise ValueError(f"unknown eigensolver {eig!r}")
        colsum = np.asarray(colsum, dtype=np.float64)
        bm = colsum / n
        n_seen = float(self.n_samples_seen_)
        n_total = n_seen + n
        if n_seen == 0:
            mc = None
        else:
            mc = math.sqrt((n_seen / n_total) * n) * (np.asarray(self.mean_) - bm)
        if on_dev:
            diag_g = G.diagonal().cpu().numpy()
            dev = G.device
            t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
            M = torch.empty_like(G)
            prev = self.components_ is not None and n_seen > 0
            # device copies stay referenced until the launch is queued (a temporary's block
            # could otherwise be handed to the next copy before the kernel reads it)
            bm_d = t(bm)
            Vt_d = t(self.components_) if prev else None
            s2_d = t(self.singular_values_ ** 2) if prev else None
            mc_d = t(mc) if mc is not None else None
            _lib.call("sh_merge_matrix", ptr(G), d, float(n), ptr(bm_d), ptr(Vt_d), ptr(s2_d),
                      self.n_components if prev else 0, ptr(mc_d), ptr(M), stream_handle(stream))
            trace_m = float(M.diagonal().sum().item())
        else:
            diag_g = np.diag(G)
            C = G - n * np.outer(bm, bm)
            C = 0.5 * (C + C.T)
            if n_seen == 0:
                M = C
            else:
                M = (self.components_.T * self.singular_values_ ** 2) @ self.components_
                M += C
                M += np.outer(mc, mc)
            trace_m = float(np.trace(M))
        new_unnorm = np.clip(diag_g - n * (bm * bm), 0.0, None)
        if n_seen == 0:
            col_mean = bm
            col_var = new_unnorm / n_total
        else:
            last_sum = np.asarray(self.mean_) * n_seen
            col_mean = (last_sum + colsum) / n_total
            ratio = n_seen / n
            upd = np.asarray(self.var_) * n_seen + new_unnorm + \
                ratio / n_total * (last_sum / ratio - colsum) ** 2
            col_var = upd / n_total
        nc = self.n_components
        w, Vt = self._eigensolve(M, nc, on_dev, eig, stream)
        w = np.clip(w, 0.0, None)
        Vt = _flip_rows(Vt)
        S = np.sqrt(w)
        rows_aug = n if n_seen == 0 else n + nc + 1
        self.n_samples_seen_ = int(n_total)
        self.components_ = Vt
        self.singular_values_ = S
        self.mean_ = col_mean
        self.var_ = col_var
        self.explained_variance_ = S ** 2 / (n_total - 1)
        tot = float(np.sum(col_var * n_total))
        self.explained_variance_ratio_ = S ** 2 / tot if tot > 0 else np.zeros_like(S)
        rank = min(rows_aug, d)
        if nc not in (rows_aug, d) and rank > nc:
            rest = max(trace_m - float(w.sum()), 0.0)
            self.noise_variance_ = rest / (n_total - 1) / (rank - nc)
        else:
            self.noise_variance_ = 0.0

    def _eigensolve(self, M, nc: int, on_dev: bool, eig: str, stream):
        d = int(M.shape[0])
        use_dev = on_dev and (eig in ("device", "lanczos")
                              or (eig == "auto" and d >= EIG_DEVICE_MIN_D))
        if eig in ("device", "lanczos") and not on_dev:
            raise ValueError(f"eig={eig!r} needs the Gram on the GPU")
        if use_dev:
            try:
                if eig == "lanczos":
                    w, Vt, info = top_eigh_device(M, nc, stream=stream)
                else:  # warm start from the previous partial fit's components
                    w, Vt, info = top_eigh_chebyshev(M, nc, stream=stream,
                                                     start=self.components_)
                self.eig_log.append(("device", info.vectors))
                return w, Vt
            except EigenNotConverged:
                if eig in ("device", "lanczos"):
                    raise
        Mh = M.cpu().numpy() if on_dev else M
        self.eig_log.append(("host", d))
        return _top_eigh(Mh, nc)  # raises LinAlgError on failure, like the SVD


def scarf_partial_fit_calls(block_rows: List[np.ndarray], dims: int):
    """Group the blocks' row lists into scarf's partial_fit calls (ann.py:215-250).

    Returns the list of calls (each a row-index array) and the held-back reservoir
    call, which is fitted last (possibly with a trailing carry-over stacked on).
    """
    calls, carry, reservoir = [], None, None
    for rows in block_rows:
        cur = rows if carry is None else np.concatenate([carry, rows])
        carry = None
        if cur.shape[0] < dims + 1:
            carry = cur
            continue
        if reservoir is None:
            reservoir = cur
            continue
        calls.append(cur)
    return calls, reservoir, carry


@dataclasses.dataclass
class PcaFit:
    loadings: np.ndarray  # f64 [d, dims]
    state: IPCAState
    mode: str


def _z_params(stats: FeatureStats, disable_scaling: bool):
    if disable_scaling:
        center = torch.zeros_like(stats.mu)
        scale = torch.ones_like(stats.sigma)
    else:
        center = stats.mu
        scale = 1.0 / stats.sigma
    return center.contiguous(), scale.contiguous()


def gram_of_rows(csr: DeviceCSR, stats: FeatureStats, params: NormParams, rows,
                 col_map=None, disable_scaling=False, sub_rows: int = 32768,
                 precision: str = "fp64", stream=None, to_host: bool = True,
                 center=None, scale=None, method: str = "auto"):
    """(Z^T Z, column sums, n) of z over ``rows`` (device int64 tensor) in FP64: host
    arrays, or with ``to_host=False`` the Gram as a device tensor (column sums on host).

    ``method="sparse"`` sums the products of the stored entries only (``sh_sparse_gram``:
    z = c + delta with c the z of a zero count, delta^T delta in exact int64 fixed point);
    ``"dense"`` densifies z (``sh_densify``) and multiplies on MFMA (``sh_gram``): FP64 z
    and FP64 MFMA with ``precision="fp64"`` (parity with the reference's FP64 PCA on every
    component), z rounded to FP32 with ``"fp32"``.  ``"auto"`` takes the sparse form for
    FP64 when the rows are at most SPARSE_GRAM_MAX_DENSITY dense.
    z = (x - center) * scale, by default the stats' mu and 1 / sigma; ``center`` /
    ``scale`` (device FP64 [d]) replace them.
    """
    require_gpu()
    d = stats.d
    ld = int(math.ceil((d + 1) / 128) * 128)  # column d carries ones -> column sums
    dev = csr.rowptr.device
    n = int(rows.shape[0])
    f64 = precision == "fp64"
    if precision not in ("fp64", "fp32"):
        raise ValueError(f"unknown Gram precision {precision!r}")
    if method not in ("auto", "sparse", "dense"):
        raise ValueError(f"unknown Gram method {method!r}")
    if method == "sparse" and not f64:
        raise ValueError("the sparse Gram is FP64 only")
    G = torch.zeros((d, d), dtype=torch.float64, device=dev)
    cs = torch.zeros(d, dtype=torch.float64, device=dev)
    c0, s0 = _z_params(stats, disable_scaling)
    center = c0 if center is None else center
    scale = s0 if scale is None else scale
    st = stream_handle(stream)
    clip = float(params.clip) if params.clip else 0.0
    if method != "dense" and f64 and d <= SPARSE_GRAM_MAX_D:
        nnz = int((csr.rowptr[rows + 1] - csr.rowptr[rows]).sum().item()) if n else 0
        if method == "sparse" or nnz <= SPARSE_GRAM_MAX_DENSITY * n * d:
            ws = workspace("gram")
            buf, nbytes = ws.get(_lib.lib().sh_sparse_gram_workspace(n, nnz, d))
            _lib.call("sh_sparse_gram", ptr(csr.vals), ptr(csr.cols), ptr(csr.rowptr),
                      ptr(rows), n, nnz, ptr(col_map), csr.n_cols, d,
                      float(params.size_factor), int(params.log_transform),
                      ptr(stats.libsize), ptr(center), ptr(scale), clip, 0, ptr(G), ptr(cs),
                      ptr(buf), nbytes, st)
            return (G.cpu().numpy() if to_host else G), cs.cpu().numpy(), n
    elif method == "sparse":
        raise ValueError(f"the sparse Gram takes at most {SPARSE_GRAM_MAX_D} features, not {d}")
    zb = min(n, sub_rows)
    Z = torch.empty((max(zb, 1), ld), dtype=torch.float64 if f64 else torch.float32, device=dev)
    ws = workspace("gram")
    for a in range(0, n, sub_rows):
        b = min(n, a + sub_rows)
        _lib.call("sh_densify", ptr(csr.vals), ptr(csr.cols), ptr(csr.rowptr), ptr(rows, a),
                  b - a, ptr(col_map), csr.n_cols, d, float(params.size_factor),
                  int(params.log_transform), ptr(stats.libsize), ptr(center), ptr(scale), clip,
                  int(f64), ptr(Z), ld, st)
        buf, nbytes = ws.get(_lib.lib().sh_gram_workspace(b - a, d, int(f64)))
        _lib.call("sh_gram", ptr(Z), int(f64), b - a, d, ld, 1, int(a > 0), ptr(G), ptr(cs),
                  ptr(buf), nbytes, st)
    return (G.cpu().numpy() if to_host else G), cs.cpu().numpy(), n


def fit_pca(csr: DeviceCSR, stats: FeatureStats, params: NormParams, dims: int,
            chunk_rows: int, use_for_pca: Optional[np.ndarray] = None, mode: str = "ipca",
            col_map=None, disable_scaling: bool = False, precision: str = "fp64",
            eig: str = "auto", stream=None, center=None, scale=None) -> PcaFit:
    """Fits dims+1 components and keeps the first ``dims`` as loadings (ann.py:212,256).

    ``eig`` selects the eigensolver of each partial fit (see ``IPCAState.merge``).
    z = (x - center) * scale; ``center`` / ``scale`` (device FP64 [d]) default to the
    stats' mu and 1 / sigma (AnnStream passes the mu / sigma it was given, as scarf's
    ``transform_z`` uses ``self.mu`` / ``self.sigma``, ``ann.py:191-192, 228-229``).
    """
    require_gpu()
    n = csr.n_rows
    dev = csr.rowptr.device
    if use_for_pca is not None:
        use_for_pca = np.asarray(use_for_pca, dtype=bool)
        if use_for_pca.shape[0] != n:
            raise ValueError("ERROR: `use_for_pca` does not have sample length as nCells")
    state = IPCAState(dims + 1)
    blocks = []
    for s, e in chunk_bounds(n, chunk_rows):
        r = np.arange(s, e, dtype=np.int64)
        if use_for_pca is not None:
            r = r[use_for_pca[s:e]]
        blocks.append(r)

    def fit_call(rows_np):
        rows = torch.from_numpy(np.ascontiguousarray(rows_np)).to(dev)
        G, cs, m = gram_of_rows(csr, stats, params, rows, col_map, disable_scaling,
                                precision=precision, stream=stream, to_host=eig == "host",
                                center=center, scale=scale)
        state.merge(m, G, cs, eig=eig, stream=stream)

    if mode == "exact":
        rows_all = np.concatenate(blocks) if blocks else np.zeros(0, np.int64)
        fit_call(rows_all)
    elif mode == "ipca":
        calls, reservoir, carry = scarf_partial_fit_calls(blocks, dims)
        pending = None
        for c in calls:
            if pending is not None:
                c = np.concatenate([pending, c])
                pending = None
            try:
                fit_call(c)
            except np.linalg.LinAlgError:
                pending = c  # carried over to the next call (ann.py:242-244)
        carry_all = [x for x in (pending, carry) if x is not None and x.shape[0]]
        last = reservoir
        if carry_all:
            last = np.concatenate(([reservoir] if reservoir is not None else []) + carry_all)
        if last is None:
            raise ValueError("no rows to fit the PCA")
        try:
            fit_call(last)
        except np.linalg.LinAlgError:
            pass
    else:
        raise ValueError(f"unknown PCA mode {mode!r}")
    if state.components_ is None:
        raise np.linalg.LinAlgError("PCA fit failed on every block")
    return PcaFit(state.components_[:-1, :].T.copy(), state, mode)


@dataclasses.dataclass
class Projector:
    """Device copies of what the sparse projection needs (reducer of ann.py:138)."""

    W: "torch.Tensor"  # f32 [d, D]
    mu: "torch.Tensor"  # f32 [d]
    inv_sigma: "torch.Tensor"  # f32 [d]
    c: "torch.Tensor"  # f32 [D]
    clip: float
    D: int
    d: int

    @classmethod
    def build(cls, loadings: np.ndarray, mu: np.ndarray, sigma: np.ndarray,
              clip: Optional[float] = None, disable_scaling: bool = False,
              fill_missing: Optional[np.ndarray] = None, offset: Optional[np.ndarray] = None,
              device="cuda") -> "Projector":
        """``fill_missing`` (bool [d]) marks features absent from the input, which scarf's
        align_features fills with 1.0 (``mapping_utils.py:208``); ``offset`` [D] is added
        to every projected row."""
        L = np.asarray(loadings, dtype=np.float64)
        d, D = L.shape
        mu = np.asarray(mu, dtype=np.float64)
        sigma = np.asarray(sigma, dtype=np.float64)
        if disable_scaling:
            mu = np.zeros(d)
            sigma = np.ones(d)
        z0 = (0.0 - mu) / sigma
        if clip:
            z0 = np.clip(z0, -clip, clip)
        zc = z0.copy()
        if fill_missing is not None:
            fm = np.asarray(fill_missing, dtype=bool)
            z1 = (1.0 - mu) / sigma
            if clip:
                z1 = np.clip(z1, -clip, clip)
            zc[fm] = z1[fm]
        c = zc @ L
        if offset is not None:
            c = c + np.asarray(offset, dtype=np.float64)
        # one host buffer, one copy: [W (d x Dp) | mu (d) | 1/sigma (d) | c (D)] in FP32,
        # W rows padded with zeros to Dp = a multiple of 4 floats (16-byte row gathers)
        Dp = -(-D // 4) * 4
        Wp = np.zeros((d, Dp), dtype=np.float32)
        Wp[:, :D] = L
        host = np.concatenate([Wp.ravel(), mu.astype(np.float32),
                               (1.0 / sigma).astype(np.float32), c.astype(np.float32)])
        dev = torch.from_numpy(host).to(device=device)
        W = dev[:d * Dp].view(d, Dp)
        o = d * Dp
        return cls(W, dev[o:o + d], dev[o + d:o + 2 * d], dev[o + 2 * d:o + 2 * d + D],
                   float(clip) if clip else 0.0, D, d)

    def project(self, csr: DeviceCSR, params: NormParams, libsize=None, col_map=None,
                row_start: int = 0, n_rows: Optional[int] = None, out=None, stream=None):
        """E [n_rows, D] FP32 for rows [row_start, row_start + n_rows) of ``csr``.

        ``libsize`` None: library size over the mapped columns, computed in-kernel.
        """
        require_gpu()
        n_rows = csr.n_rows - row_start if n_rows is None else int(n_rows)
        if out is None:
            out = torch.empty((n_rows, self.D), dtype=torch.float32, device=csr.rowptr.device)
        if n_rows == 0:
            return out
        _lib.call("sh_project", ptr(csr.vals), ptr(csr.cols), ptr(csr.rowptr, row_start), n_rows,
                  ptr(col_map), csr.n_cols, self.d, float(params.size_factor),
                  int(params.log_transform),
                  ptr(libsize, row_start) if libsize is not None else None, ptr(self.mu),
                  ptr(self.inv_sigma), self.clip, ptr(self.W), self.W.stride(0), self.D,
                  ptr(self.c), ptr(out), out.stride(0), stream_handle(stream))
        return out
