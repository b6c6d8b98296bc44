"""The graph-building pipeline (scarf ``DataStore.make_graph``) on resident CSR counts.

Stages, each a HIP kernel sequence on the current stream:
  1. library size, log1p normalisation and per-chunk FP64 feature sums, merged to
     mu / sigma with clean_array      (graph_datastore.py:767-796, assay.py:780-828)
  2. PCA: dims + 1 components, FP64 MFMA Gram per partial_fit call + FP64 host merge
     and eigensolve; scarf's block order ("ipca") or one exact fit ("exact")
                                        (ann.py:207-256)
  3. sparse z-score projection E = z W (ann.py:129-138)
  4. self kNN: k + 1 neighbours, screen on MFMA, exact FP64 re-rank, fix_knn_query
     self repair, recall %               (knn_utils.py:23-76, ann.py:31-52)
  5. smooth_knn_dist / membership weights, zero -> global min (knn_utils.py:89-159)

Outputs mirror what make_graph stores: mu, sigma, loadings ("reduction"),
indices / distances ("knn__k"), edges / weights ("graph__lc__bw").
"""
from __future__ import annotations

import dataclasses
import time
from typing import Dict, Optional

import numpy as np

from .device import DeviceCSR, col_map_tensor, require_gpu
from .knn import GpuIndex, pack_refs, search
from .normalize import FeatureStats, HostCSR, NormParams, feature_stats, upload_with_stats
from .pca import PcaFit, Projector, fit_pca
from .smooth import edges_of, smooth_knn

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


@dataclasses.dataclass
class GraphParams:
    """make_graph's knobs with scarf's defaults (graph_datastore.py:139-149, 272, 298-303)."""

    dims: int = 11
    k: int = 11
    batch_size: int = 1000  # rows per chunk (scarf: the raw zarr chunk)
    size_factor: float = 1000.0
    log_transform: bool = True
    renormalize_subset: bool = True
    clip: Optional[float] = None
    feat_scaling: bool = True
    local_connectivity: float = 1.0
    bandwidth: float = 1.5
    pca_mode: str = "ipca"  # "ipca" (scarf's block order) or "exact"
    gram_precision: str = "fp64"
    pca_eig: str = "auto"  # "auto" | "device" (GPU Chebyshev) | "lanczos" | "host" (LAPACK)
    screen: str = "bf16"  # "fp32" = certified exact, "bf16" = screened
    pool: int = 64

    def norm(self) -> NormParams:
        return NormParams(self.size_factor, self.log_transform, self.renormalize_subset, self.clip)


@dataclasses.dataclass
class GraphResult:
    mu: np.ndarray
    sigma: np.ndarray
    loadings: np.ndarray
    pca: PcaFit
    stats: FeatureStats
    embedding: "torch.Tensor"
    indices: "torch.Tensor"
    distances: "torch.Tensor"
    weights: "torch.Tensor"
    recall: float
    n_mis: int
    timings_ms: Dict[str, float]
    index: GpuIndex
    n_refined: int = 0  # FP8 screen: queries screened a second time (SH_FLAG_THIN)

    def edges(self) -> np.ndarray:
        n, k = self.indices.shape
        return edges_of(n, k, self.indices.cpu().numpy())


class StageTimer:
    """HIP-event timing of stages on the current stream (plus host wall time)."""

    def __init__(self, enabled=True):
        self.enabled = enabled
        self.marks = []

    def mark(self, name):
        if not self.enabled:
            return
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.marks.append((name, ev, time.perf_counter()))

    def result(self) -> Dict[str, float]:
        if not self.enabled or len(self.marks) < 2:
            return {}
        torch.cuda.synchronize()
        out = {}
        for (a, ea, _), (b, eb, _) in zip(self.marks[:-1], self.marks[1:]):
            out[b] = ea.elapsed_time(eb)
        out["total"] = self.marks[0][1].elapsed_time(self.marks[-1][1])
        out["wall_total"] = (self.marks[-1][2] - self.marks[0][2]) * 1e3
        return out


def build_graph(csr: DeviceCSR, params: GraphParams, feat_idx=None,
                use_for_pca: Optional[np.ndarray] = None,
                n_counts: Optional["torch.Tensor"] = None,
                loadings: Optional[np.ndarray] = None, timer: Optional[StageTimer] = None,
                stream=None, stats: Optional[FeatureStats] = None) -> GraphResult:
    """The whole path on resident counts.  ``stats`` (from :func:`build_graph_streamed`'s
    upload) skips the statistics pass."""
    require_gpu()
    timer = timer or StageTimer(False)
    p = params.norm()
    n = csr.n_rows
    k = min(params.k, n - 1)
    cm, d = col_map_tensor(csr.n_cols, feat_idx, device=csr.rowptr.device)
    if stats is None:
        timer.mark("start")
        stats = feature_stats(csr, params.batch_size, p, col_map=cm, d=d, n_counts=n_counts,
                              stream=stream)
        timer.mark("normalise_stats")
    mu = stats.mu.cpu().numpy()
    sigma = stats.sigma.cpu().numpy()
    dims = min(params.dims, n, max(1, params.batch_size - 1))
    if loadings is None:
        fit = fit_pca(csr, stats, p, dims, params.batch_size, use_for_pca=use_for_pca,
                      mode=params.pca_mode, col_map=cm,
                      disable_scaling=not params.feat_scaling,
                      precision=params.gram_precision, eig=params.pca_eig, stream=stream)
    else:
        fit = PcaFit(np.asarray(loadings, dtype=np.float64), None, "given")
    timer.mark("pca_fit")
    proj = Projector.build(fit.loadings, mu, sigma, clip=params.clip,
                           disable_scaling=not params.feat_scaling, device=csr.rowptr.device)
    E = proj.project(csr, p, libsize=stats.libsize, col_map=cm, stream=stream)
    timer.mark("project")
    index = GpuIndex("l2", E.shape[1], screen=params.screen, pool=params.pool,
                     device=csr.rowptr.device)
    index.init_index(n)
    index.add_items(E)
    refs = index.packed()
    idx, dist, flags = search(E, E, refs, k, params.pool, self_offset=0, stream=stream)
    timer.mark("knn")
    _, _, w = smooth_knn(dist, idx, params.batch_size, params.local_connectivity,
                         params.bandwidth, stream=stream)
    timer.mark("smooth")
    # fix_knn_query's n_mis: rows whose first neighbour (of k + 1) was not the row itself
    n_mis = int(((flags & 4) > 0).sum().item())
    n_refined = int(((flags & 8) > 0).sum().item())
    recall = 100.0 * (n - n_mis) / max(n, 1)
    return GraphResult(mu, sigma, fit.loadings, fit, stats, E, idx, dist, w, recall, n_mis,
                       timer.result(), index, n_refined)


def build_graph_streamed(host: HostCSR, params: GraphParams, feat_idx=None,
                         timer: Optional[StageTimer] = None, stream=None,
                         copy_stream=None) -> GraphResult:
    """The whole path from counts in pinned host memory: the CSR streams to HBM in
    ``params.batch_size``-row chunks while each landed chunk's statistics are computed
    (``upload_with_stats``), then the rest runs on the resident copy."""
    require_gpu()
    timer = timer or StageTimer(False)
    cm, d = col_map_tensor(host.n_cols, feat_idx)
    timer.mark("start")
    csr, stats = upload_with_stats(host, params.batch_size, params.norm(), col_map=cm, d=d,
                                   copy_stream=copy_stream, stream=stream)
    timer.mark("upload_stats")
    return build_graph(csr, params, feat_idx=feat_idx, timer=timer, stream=stream, stats=stats)
