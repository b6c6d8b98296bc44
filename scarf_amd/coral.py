"""CORAL alignment for mapping (scarf ``mapping_utils.py:15-95``, ``run_coral=True`` in
``mapping_datastore.py:160-175``).

scarf standardises the aligned query data by its own feature means and standard
deviations (``clean_array``, ddof 0), the reference data by its own, takes both feature
covariances (ddof 1, plus the identity), and rewrites the query as
``data_coral = s @ C_s^{-1/2} C_t^{1/2}``.  run_mapping turns feature scaling off when CORAL
runs (``mapping_datastore.py:141-142``), so the reference AnnStream reduces ``data_coral``
with its loadings alone, ``data_coral @ L`` (``graph_datastore.py:886``, ``ann.py:126-127``).

Here nothing of size cells x features is formed:

* both covariances come from FP64 Grams of the standardised rows (``gram_of_rows`` with
  the query's or the reference's centering: the sparse form with exact integer sums for
  sparse counts, dense MFMA otherwise), turned into C = cov + I on the device
  (``sh_cov_from_gram``); the reference's is computed once per ``AnnStream``;
* C_s^{-1/2} and C_t^{1/2} come from the coupled Newton-Schulz iteration on FP64 GEMMs
  (``sh_spd_sqrt``) instead of scipy's ``fractional_matrix_power``;
* the whole map x -> (s @ A) @ L is linear in x, so it folds into the sparse projection
  kernel: weights B = A L (an FP64 GEMM) over s = (x - sm) / sd.
"""
from __future__ import annotations

import dataclasses
import logging

import numpy as np

from . import _lib
from .device import ptr, require_gpu, stream_handle, workspace
from .normalize import feature_stats
from .pca import Projector, gram_of_rows

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

logger = logging.getLogger("scarf_amd")

__all__ = ["spd_sqrt", "covariance_plus_identity", "coral_projector", "CoralInfo"]


@dataclasses.dataclass
class CoralInfo:
    iters_source: int
    resid_source: float
    iters_target: int
    resid_target: float


def spd_sqrt(C: "torch.Tensor", sqrt: bool = True, inv_sqrt: bool = True, tol: float = 1e-13,
             max_iter: int = 100):
    """(C^{1/2} or None, C^{-1/2} or None, iterations, residual) of an SPD FP64 device
    matrix; raises if the iteration does not reach 1e-8 * sqrt(d)."""
    require_gpu()
    d = int(C.shape[0])
    C = C.contiguous()
    out_s = torch.empty_like(C) if sqrt else None
    out_i = torch.empty_like(C) if inv_sqrt else None
    it = np.zeros(1, dtype=np.int32)
    res = np.zeros(1, dtype=np.float64)
    ws, nbytes = workspace("spd_sqrt").get(_lib.lib().sh_spd_sqrt_workspace(d))
    _lib.call("sh_spd_sqrt", ptr(C), d, ptr(out_s), ptr(out_i), int(max_iter), float(tol),
              it.ctypes.data_as(_lib.ctypes.c_void_p), res.ctypes.data_as(_lib.ctypes.c_void_p),
              ptr(ws), nbytes, stream_handle())
    if not res[0] <= 1e-8 * np.sqrt(d):
        raise RuntimeError(f"matrix square root did not converge: residual {res[0]:.3g} after "
                           f"{int(it[0])} iterations")
    return out_s, out_i, int(it[0]), float(res[0])


def covariance_plus_identity(csr, stats, params, center, scale, col_map=None):
    """cov(z) + I (ddof 1) of z = (x - center) * scale over every row of ``csr``, FP64 on the
    device (``mapping_utils.py:15-19``)."""
    dev = csr.rowptr.device
    rows = torch.arange(csr.n_rows, dtype=torch.int64, device=dev)
    G, cs, n = gram_of_rows(csr, stats, params, rows, col_map=col_map, to_host=False,
                            center=center, scale=scale)
    d = stats.d
    cs_d = torch.from_numpy(np.ascontiguousarray(cs, dtype=np.float64)).to(dev)
    C = torch.empty((d, d), dtype=torch.float64, device=dev)
    _lib.call("sh_cov_from_gram", ptr(G), G.stride(0), ptr(cs_d), n, d, 1.0, ptr(C),
              stream_handle())
    return C


def _dgemm(A, B):
    m, k = A.shape
    n = B.shape[1]
    out = torch.empty((m, n), dtype=torch.float64, device=A.device)
    _lib.call("sh_dgemm", 0, m, n, k, 1.0, ptr(A), A.stride(0), ptr(B), B.stride(0), 0.0,
              ptr(out), out.stride(0), stream_handle())
    return out


def reference_covariance(ann):
    """C_t = cov(t) + I of the reference standardised by its own clean mean / std
    (``mapping_utils.py:80-90``), cached on the AnnStream."""
    T = getattr(ann, "_coral_target_cov", None)
    if T is None:
        data = ann.data
        st = data.stats()
        T = covariance_plus_identity(data.counts, st, data.norm, st.mu.contiguous(),
                                     (1.0 / st.sigma).contiguous(), col_map=data._cm)
        ann._coral_target_cov = T
    return T


def coral_projector(ann, tcsr, col_map, missing, norm, n_counts=None):
    """(Projector of the CORAL-corrected query onto the reference embedding, CoralInfo)."""
    require_gpu()
    if norm.clip:
        raise NotImplementedError("CORAL with clipped z-scores is not defined by scarf")
    dev = tcsr.rowptr.device
    d = int(np.asarray(ann.loadings).shape[0])
    missing = np.asarray(missing, dtype=bool)
    # query standardisation: its own clean mean / std over the aligned data, whose missing
    # features are constant 1.0 (mean 1, std 0 -> clean 1: they standardise to 0)
    tst = feature_stats(tcsr, 1000, norm, col_map=col_map, d=d, n_counts=n_counts)
    sm = tst.mu.cpu().numpy().copy()
    sd = tst.sigma.cpu().numpy().copy()
    sm[missing] = 1.0
    sd[missing] = 1.0
    center = torch.from_numpy(np.where(missing, 0.0, sm)).to(dev)
    scale = torch.from_numpy(np.where(missing, 0.0, 1.0 / sd)).to(dev)
    S = covariance_plus_identity(tcsr, tst, norm, center, scale, col_map=col_map)
    T = reference_covariance(ann)
    _, S_mh, its, rs = spd_sqrt(S, sqrt=False, inv_sqrt=True)
    T_h, _, itt, rt = spd_sqrt(T, sqrt=True, inv_sqrt=False)
    A = _dgemm(S_mh, T_h)
    # scarf runs CORAL with feature scaling off (mapping_datastore.py:141-142): the
    # reference AnnStream then reduces data_coral @ loadings with no z-scoring
    # (graph_datastore.py:886, ann.py:126-127), whatever the graph was built with
    L = np.asarray(ann.loadings, dtype=np.float64)
    if not ann._disable_scaling:
        logger.warning("run_coral: scarf maps CORAL-corrected data with feature scaling "
                       "off; this reference AnnStream was fitted with feat_scaling=True, "
                       "so its loadings are applied to the unscaled corrected data")
    B = _dgemm(A, torch.from_numpy(np.ascontiguousarray(L)).to(dev)).cpu().numpy()
    proj = Projector.build(B, sm, sd, clip=None, fill_missing=missing)
    return proj, CoralInfo(its, rs, itt, rt)
