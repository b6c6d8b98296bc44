"""Drop-in for scarf's ``AnnStream`` (``ann.py:55-346``) on GPU-resident CSR counts.

scarf hands ``AnnStream`` a dask array of normalised values chunked by ``batch_size``
rows.  Here ``data`` is a :class:`NormedCSR`: the raw CSR counts resident in HBM plus
the normalisation of ``RNAassay.normed`` (size factor, log1p, subset library size,
feature subset), exposing the same block interface (``shape``, ``chunksize``,
``numblocks``, ``blocks``).  Normalised values are recomputed inside each kernel from
the counts and are never materialised.

Attributes read by scarf's callers are provided with the reference's names:
``data, k, nCells, nFeats, dims, loadings, mu, sigma, method, reducer,
transform_ann, transform_z, iter_blocks, annIdx, kmeans, clusterLabels,
harmonizedData, _pca.explained_variance_ratio_`` (``graph_datastore.py:889-1015``,
``mapping_datastore.py:143-208``, ``metrics.py:290-318``).
"""
from __future__ import annotations

import dataclasses
import logging
from typing import Optional

import numpy as np

from . import _lib
from .device import DeviceCSR, col_map_tensor, ptr, require_gpu, stream_handle
from .kmeans import MiniBatchKMeans
from .knn import GpuIndex, fix_knn_query_device
from .normalize import FeatureStats, NormParams, chunk_bounds, feature_stats
from .pca import PcaFit, Projector, _z_params, fit_pca

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

logger = logging.getLogger("scarf_amd")


def _dense_code(X) -> int:
    return _lib.SH_F64 if X.dtype == torch.float64 else _lib.SH_F32

__all__ = ["AnnStream", "NormedCSR", "instantiate_knn_index", "fix_knn_query"]


@dataclasses.dataclass
class NormedCSR:
    """Normalised view of resident CSR counts, blocked by ``batch_size`` rows."""

    counts: DeviceCSR
    norm: NormParams
    batch_size: int = 1000
    feat_idx: Optional[np.ndarray] = None
    n_counts: Optional["torch.Tensor"] = None  # per-cell totals when renormalize_subset=False
    row_start: int = 0
    row_end: Optional[int] = None
    _stats: Optional[FeatureStats] = None
    _cm: object = None
    _d: Optional[int] = None

    def __post_init__(self):
        if self.row_end is None:
            self.row_end = self.counts.n_rows
        if self._d is None:
            self._cm, self._d = col_map_tensor(self.counts.n_cols, self.feat_idx,
                                               device=self.counts.rowptr.device)

    @property
    def shape(self):
        return (self.row_end - self.row_start, self._d)

    @property
    def chunksize(self):
        return (self.batch_size, self._d)

    @property
    def numblocks(self):
        return (len(self._bounds()), 1)

    def _bounds(self):
        n = self.row_end - self.row_start
        return [(self.row_start + s, self.row_start + e) for s, e in chunk_bounds(n, self.batch_size)
                if e > s]

    @property
    def blocks(self):
        return [self.view(s, e) for s, e in self._bounds()]

    def view(self, s: int, e: int) -> "NormedCSR":
        self.stats()  # library sizes and mu/sigma are shared by every view
        return dataclasses.replace(self, row_start=s, row_end=e)

    def stats(self) -> FeatureStats:
        """Library sizes and mu/sigma over the whole matrix (computed once)."""
        if self._stats is None:
            self._stats = feature_stats(self.counts, self.batch_size, self.norm, col_map=self._cm,
                                        d=self._d, n_counts=self.n_counts)
        return self._stats

    def dense(self, center=None, scale=None) -> np.ndarray:
        """FP64 dense normalised rows (x, or (x - center) * scale) of this view."""
        from . import _lib
        st = self.stats()
        n, d = self.shape
        ld = int(np.ceil(d / 4) * 4)
        dev = self.counts.rowptr.device
        out = torch.empty((max(n, 1), ld), dtype=torch.float64, device=dev)
        c = torch.zeros(d, dtype=torch.float64, device=dev) if center is None else center
        sc = torch.ones(d, dtype=torch.float64, device=dev) if scale is None else scale
        rows = torch.arange(self.row_start, self.row_end, device=dev, dtype=torch.int64)
        c, sc = c.contiguous(), sc.contiguous()  # referenced until the launch is queued
        _lib.call("sh_densify", ptr(self.counts.vals), ptr(self.counts.cols),
                  ptr(self.counts.rowptr), ptr(rows), n, ptr(self._cm), self.counts.n_cols, d,
                  float(self.norm.size_factor), int(self.norm.log_transform), ptr(st.libsize),
                  ptr(c), ptr(sc),
                  float(self.norm.clip) if self.norm.clip else 0.0, 1, ptr(out), ld,
                  stream_handle())
        return out[:n, :d].cpu().numpy()


def instantiate_knn_index(space, dim, max_elements, ef_construction, M, random_seed, ef,
                          num_threads, screen: str = "bf16", pool: int = 64):
    """``ann.py:14-28`` with the GPU index in place of hnswlib."""
    idx = GpuIndex(space=space, dim=dim, screen=screen, pool=pool)
    idx.init_index(max_elements=max_elements, ef_construction=ef_construction, M=M,
                   random_seed=random_seed)
    idx.set_ef(ef)
    idx.set_num_threads(num_threads)
    return idx


def fix_knn_query(indices, distances, ref_idx):
    """``ann.py:31-52`` on the GPU: returns (indices[:, k], distances[:, k], n_mis)."""
    require_gpu()
    host = not isinstance(indices, torch.Tensor)
    i = torch.as_tensor(np.asarray(indices).astype(np.int64) if host else indices,
                        dtype=torch.int64).cuda()
    d = torch.as_tensor(np.asarray(distances, dtype=np.float32) if host else distances,
                        dtype=torch.float32).cuda()
    r = torch.as_tensor(np.asarray(ref_idx).astype(np.int64) if host else ref_idx,
                        dtype=torch.int64).cuda()
    oi, od, nm = fix_knn_query_device(i, d, r)
    if host:
        return oi.cpu().numpy().astype(np.asarray(indices).dtype), od.cpu().numpy(), nm
    return oi, od, nm


class AnnStream:
    """scarf ``AnnStream`` with the PCA fit, projection and kNN index on the GPU."""

    def __init__(self, data: NormedCSR, k: int, n_cluster: int, reduction_method: str,
                 dims: int, loadings, use_for_pca, mu, sigma, ann_metric: str, ann_efc: int,
                 ann_ef: int, ann_m: int, nthreads: int, ann_parallel: bool, rand_state: int,
                 do_kmeans_fit: bool, disable_scaling: bool, ann_idx, lsi_skip_first: bool,
                 lsi_params: dict, harmonize: bool, harmonized_data=None, batches=None,
                 screen: str = "bf16", pool: int = 64, pca_mode: str = "ipca",
                 gram_precision: str = "fp64"):
        require_gpu()
        if not isinstance(data, NormedCSR):
            raise TypeError("AnnStream data must be a NormedCSR (GPU-resident counts)")
        self.data = data
        self.k = k
        if self.k >= self.data.shape[0]:
            self.k = self.data.shape[0] - 1
        self.nClusters = max(n_cluster, 2)
        self.dims = dims
        self.loadings = loadings
        if self.dims is None and self.loadings is None:
            raise ValueError("ERROR: Provide either value for atleast one: 'dims' or 'loadings'")
        self.annMetric = ann_metric
        self.annEfc = ann_efc
        self.annEf = ann_ef
        self.annM = ann_m
        self.nthreads = nthreads
        self.annThreads = self.nthreads if ann_parallel else 1
        self.randState = rand_state
        self.batchSize = self._handle_batch_size()
        self.method = reduction_method
        self.nCells, self.nFeats = self.data.shape
        self.clusterLabels = np.repeat(-1, self.nCells)
        self.harmonize = harmonize
        self.harmonizedData = harmonized_data
        self.batches = batches
        self.screen = screen
        self.pool = pool
        self._pca = None
        self._disable_scaling = disable_scaling
        if harmonize:
            raise NotImplementedError("harmonize=True (Harmony batch correction) is outside "
                                      "this build's scope")
        disable_reduction = self.dims < 1
        st = self.data.stats()
        if self.method == "pca":
            if mu is None or np.size(mu) != self.nFeats:
                mu = st.mu.cpu().numpy()
            if sigma is None or np.size(sigma) != self.nFeats:
                sigma = st.sigma.cpu().numpy()
            self.mu, self.sigma = np.asarray(mu, dtype=np.float64), np.asarray(sigma, dtype=np.float64)
            if self.loadings is None or len(self.loadings) == 0:
                if use_for_pca is not None and len(use_for_pca) != self.nCells:
                    raise ValueError("ERROR: `use_for_pca` does not have sample length as nCells")
                if not disable_reduction:
                    self._fit_pca(disable_scaling, use_for_pca, pca_mode, gram_precision)
            else:
                self.loadings = np.asarray(self.loadings, dtype=np.float64)
                self.dims = self.loadings.shape[1]
        elif self.method == "custom":
            self.mu = np.zeros(self.nFeats)
            self.sigma = np.ones(self.nFeats)
            disable_scaling = True
            self._disable_scaling = True
            if self.loadings is None or len(self.loadings) == 0:
                logger.warning("No loadings provided for manual dimension reduction")
            else:
                self.loadings = np.asarray(self.loadings, dtype=np.float64)
                self.dims = self.loadings.shape[1]
        elif self.method == "lsi":
            raise NotImplementedError("LSI (ATAC, gensim) is outside this build's scope")
        else:
            raise ValueError(f"ERROR: Unknown reduction method: {self.method}")
        if disable_reduction:
            raise NotImplementedError("dims < 1 (no reduction) is outside this build's scope")
        self._proj = Projector.build(self.loadings, self.mu, self.sigma, clip=self.data.norm.clip,
                                     disable_scaling=self._disable_scaling)
        if ann_idx is None:
            self.annIdx = self._fit_ann()
        else:
            self.annIdx = ann_idx
            self.annIdx.set_ef(self.annEf)
            self.annIdx.set_num_threads(self.annThreads)
        self.kmeans = self._fit_kmeans(do_kmeans_fit)

    # -- ann.py:173-185 ------------------------------------------------------------------
    def _handle_batch_size(self):
        if self.dims > self.data.shape[0]:
            self.dims = self.data.shape[0]
        batch_size = self.data.chunksize[0]
        if self.dims >= batch_size:
            self.dims = batch_size - 1
            logger.info(f"Number of PCA/LSI components reduced to batch size of {batch_size}")
        if self.nClusters > batch_size:
            self.nClusters = batch_size
        return batch_size

    def iter_blocks(self, msg: str = ""):
        for b in self.data.blocks:
            yield b

    def transform_z(self, a):
        """z = (a - mu) / sigma of a block, ``ann.py:191-192``: a NormedCSR view (its rows
        densified from the counts), a dense numpy block of normalised values (FP64 numpy
        out) or a dense device tensor (FP64 tensor out)."""
        if isinstance(a, NormedCSR):
            c = torch.from_numpy(np.ascontiguousarray(self.mu)).cuda()
            s = torch.from_numpy(np.ascontiguousarray(1.0 / self.sigma)).cuda()
            return a.dense(c, s)
        X, host = self._dense_block(a)
        ops = self._dense_ops()
        z = torch.empty((max(X.shape[0], 1), X.shape[1]), dtype=torch.float64, device=X.device)
        _lib.call("sh_dense_z", ptr(X), _dense_code(X), X.shape[0], X.shape[1], X.stride(0),
                  ptr(ops["mu"]), ptr(ops["sigma"]), ops["clip"], ptr(z), z.stride(0),
                  stream_handle())
        z = z[:X.shape[0]]
        return z.cpu().numpy() if host else z

    def reducer_tensor(self, a) -> "torch.Tensor":
        """Embedding of a block on the device: FP32 for a NormedCSR view (sparse
        projection), FP64 for a dense block (``sh_dense_project``)."""
        if not isinstance(a, NormedCSR):
            X, _ = self._dense_block(a)
            ops = self._dense_ops()
            D = int(ops["W"].shape[1])
            out = torch.empty((max(X.shape[0], 1), D), dtype=torch.float64, device=X.device)
            _lib.call("sh_dense_project", ptr(X), _dense_code(X), X.shape[0], X.shape[1],
                      X.stride(0), ptr(ops["mu"]) if ops["scaled"] else None,
                      ptr(ops["sigma"]) if ops["scaled"] else None,
                      ops["clip"] if ops["scaled"] else 0.0, ptr(ops["W"]), D, ptr(out),
                      out.stride(0), stream_handle())
            return out[:X.shape[0]]
        st = a.stats()
        E = self._proj.project(a.counts, a.norm, libsize=st.libsize, col_map=a._cm,
                               row_start=a.row_start, n_rows=a.row_end - a.row_start)
        return E

    def reducer(self, a):
        """E = transform_z(a) @ loadings (or a @ loadings without scaling) of a block,
        ``ann.py:119-138``: FP64 numpy for a NormedCSR view or a numpy block, an FP64
        device tensor for a dense device tensor."""
        if isinstance(a, NormedCSR):
            return self.reducer_tensor(a).cpu().numpy().astype(np.float64)
        E = self.reducer_tensor(a)
        return E if isinstance(a, torch.Tensor) else E.cpu().numpy()

    def _dense_block(self, a):
        """(2-D FP32/FP64 device tensor with unit column stride, came-from-host flag)."""
        host = not isinstance(a, torch.Tensor)
        if host:
            arr = np.asarray(a)
            if arr.ndim == 1:
                arr = arr[None, :]
            if arr.dtype != np.float32:
                arr = arr.astype(np.float64, copy=False)
            X = torch.from_numpy(np.ascontiguousarray(arr)).cuda()
        else:
            X = a if a.dim() == 2 else a.reshape(1, -1)
            if X.dtype not in (torch.float32, torch.float64):
                X = X.to(torch.float64)
            X = X.cuda()
            if X.stride(1) != 1:
                X = X.contiguous()
        if X.shape[1] != self.nFeats:
            raise ValueError(f"block has {X.shape[1]} features, the reference has {self.nFeats}")
        return X, host

    def _dense_ops(self):
        """Device FP64 loadings, mu and sigma for dense blocks, rebuilt when the host
        attributes change (scarf's mapping replaces mu / sigma, mapping_datastore.py:176-190)."""
        mu = np.asarray(self.mu, dtype=np.float64)
        sg = np.asarray(self.sigma, dtype=np.float64)
        L = np.asarray(self.loadings, dtype=np.float64)
        c = getattr(self, "_dense_cache", None)
        if c is not None and np.array_equal(c["h_mu"], mu) and np.array_equal(c["h_sigma"], sg) \
                and c["h_W"].shape == L.shape and np.array_equal(c["h_W"], L):
            return c
        dev = self.data.counts.rowptr.device
        t = lambda v: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev)
        scaled = self.method == "pca" and not self._disable_scaling
        clip = float(self.data.norm.clip) if self.data.norm.clip else 0.0
        c = {"h_mu": mu.copy(), "h_sigma": sg.copy(), "h_W": L.copy(), "mu": t(mu),
             "sigma": t(sg), "W": t(L), "scaled": scaled, "clip": clip}
        self._dense_cache = c
        return c

    def transform_ann(self, a, k: int = None, self_indices=None):
        """``ann.py:194-205``: k neighbours, or k+1 then the self repair."""
        if k is None:
            k = self.k
        Q = a if isinstance(a, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(np.asarray(a), dtype=np.float32)).cuda()
        if self_indices is None:
            idx, dist, _ = self.annIdx.knn_query_tensor(Q.float().contiguous(), k)
            return idx.cpu().numpy().astype(np.uint64), dist.cpu().numpy()
        si = torch.from_numpy(np.asarray(self_indices).astype(np.int64)).cuda()
        idx, dist, flags = self.annIdx.knn_query_tensor(Q.float().contiguous(), k, self_idx=si)
        n_mis = int(((flags & 4) > 0).sum().item())
        return idx.cpu().numpy().astype(np.uint64), dist.cpu().numpy(), n_mis

    # -- ann.py:207-256 ------------------------------------------------------------------
    def _fit_pca(self, disable_scaling, use_for_pca, pca_mode="ipca", gram_precision="fp64"):
        st = self.data.stats()
        center = scale = None
        if not disable_scaling:  # z with the mu / sigma this object holds (ann.py:228-229)
            dev = self.data.counts.rowptr.device
            center = torch.from_numpy(np.ascontiguousarray(self.mu, dtype=np.float64)).to(dev)
            scale = torch.from_numpy(
                np.ascontiguousarray(1.0 / np.asarray(self.sigma, dtype=np.float64))).to(dev)
        fit = fit_pca(self.data.counts, st, self.data.norm, self.dims, self.batchSize,
                      use_for_pca=None if use_for_pca is None else np.asarray(use_for_pca, bool),
                      mode=pca_mode, col_map=self.data._cm, disable_scaling=disable_scaling,
                      precision=gram_precision, center=center, scale=scale)
        self._pca = fit.state
        self.loadings = fit.loadings

    # -- ann.py:328-346 ------------------------------------------------------------------
    def _fit_kmeans(self, do_ann_fit):
        """MiniBatchKMeans seed partitions: partial_fit on every block's embedding, then
        predict on every block.  The embedding stays on the device; labels are int32 like
        scikit-learn's.  ``labels_`` is not computed per step (scarf never reads it)."""
        if do_ann_fit is False:
            return None
        km = MiniBatchKMeans(n_clusters=self.nClusters, random_state=self.randState,
                             batch_size=self.batchSize, n_init=3, compute_labels=False)
        for b in self.iter_blocks(msg="Fitting kmeans"):
            km.partial_fit(self.reducer_tensor(b))
        labels = [km.predict_tensor(self.reducer_tensor(b))
                  for b in self.iter_blocks(msg="Estimating seed partitions")]
        self.clusterLabels = torch.cat(labels).cpu().numpy()
        return km

    # -- ann.py:290-326 ------------------------------------------------------------------
    def _fit_ann(self):
        idx = instantiate_knn_index(self.annMetric, self.dims, self.nCells, self.annEfc,
                                    self.annM, self.randState, self.annEf, self.annThreads,
                                    screen=self.screen, pool=self.pool)
        for b in self.iter_blocks(msg="Fitting ANN"):
            idx.add_items(self.reducer_tensor(b))
        return idx
