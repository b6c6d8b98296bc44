"""Drop-ins for ``DataStore.make_graph``, ``run_mapping`` and ``get_mapping_score``.

The reference keeps every result in a zarr hierarchy whose group names carry the
parameters (``graph_datastore.py:558-575``, ``738-745``).  These functions take the
counts as CSR (host ``scarf_amd.synth.CSR`` or ``DeviceCSR``) and write the same
datasets under the same paths into ``store`` (any mutable mapping; a dict by default):

  {assay}/normed__{cell_key}__{feat_key}/mu, /sigma
  .../reduction__pca__{dims}__{pca_cell_key}/reduction                (loadings)
  .../ann__{metric}__{efc}__{ef}__{M}__{seed}/knn__{k}/indices, /distances
  .../knn__{k}/graph__{lc}__{bw}/edges, /weights
  {assay}/projections/{target_name}/indices, /distances
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence

import numpy as np

from .ann import AnnStream, NormedCSR
from .device import DeviceCSR, col_map_from_positions, require_gpu
from .knn_utils import self_query_knn, smoothen_dists
from .mapping import align_col_map, mapping_scores
from .normalize import NormParams, feature_stats
from .pca import Projector

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

DEFAULTS = {"dims": 11, "k": 11, "ann_metric": "l2", "rand_state": 4466,
            "n_centroids": 1000, "local_connectivity": 1.0, "bandwidth": 1.5}


def _device_csr(counts) -> DeviceCSR:
    if isinstance(counts, DeviceCSR):
        return counts
    return DeviceCSR.from_host(counts)


def graph_locations(assay, cell_key, feat_key, dims, pca_cell_key, metric, efc, ef, m, seed,
                    k, lc, bw):
    normed = f"{assay}/normed__{cell_key}__{feat_key}"
    reduction = f"{normed}/reduction__pca__{dims}__{pca_cell_key}"
    ann = f"{reduction}/ann__{metric}__{efc}__{ef}__{m}__{seed}"
    knn = f"{ann}/knn__{k}"
    graph = f"{knn}/graph__{lc}__{bw}"
    return normed, reduction, ann, knn, graph


def make_graph(counts, feat_idx=None, *, assay: str = "RNA", cell_key: str = "I",
               feat_key: str = "hvgs", pca_cell_key: Optional[str] = None,
               use_for_pca: Optional[np.ndarray] = None, dims: Optional[int] = None,
               k: Optional[int] = None, ann_metric: str = "l2", ann_efc: Optional[int] = None,
               ann_ef: Optional[int] = None, ann_m: Optional[int] = None,
               rand_state: int = 4466, batch_size: int = 1000, size_factor: float = 1000.0,
               log_transform: bool = True, renormalize_subset: bool = True,
               n_counts=None, local_connectivity: float = 1.0, bandwidth: float = 1.5,
               feat_scaling: bool = True, custom_loadings: Optional[np.ndarray] = None,
               clip: Optional[float] = None, return_ann_object: bool = False,
               screen: str = "bf16", pool: int = 64, pca_mode: str = "ipca",
               n_centroids: Optional[int] = None, store: Optional[Dict] = None):
    """``GraphDataStore.make_graph`` (``graph_datastore.py:513-1020``) on CSR counts.

    Parameter defaults follow ``_set_graph_params`` (``graph_datastore.py:139-149``,
    ``272``, ``298-303``).  Returns the AnnStream if ``return_ann_object`` else the store.
    """
    require_gpu()
    store = {} if store is None else store
    if ann_metric != "l2":
        raise ValueError("only ann_metric='l2' is supported")
    dims = DEFAULTS["dims"] if dims is None else int(dims)
    method = "pca"
    if custom_loadings is not None:
        method = "custom"
        dims = custom_loadings.shape[1]
    k = DEFAULTS["k"] if k is None else int(k)
    ann_m = min(max(48, int(dims * 1.5)), 64) if ann_m is None else int(ann_m)
    ann_ef = min(100, max(k * 3, 50)) if ann_ef is None else int(ann_ef)
    ann_efc = min(100, max(k * 3, 50)) if ann_efc is None else int(ann_efc)
    pca_cell_key = cell_key if pca_cell_key is None else pca_cell_key
    csr = _device_csr(counts)
    norm = NormParams(size_factor, log_transform, renormalize_subset, clip)
    nc = None
    if n_counts is not None:
        nc = torch.as_tensor(np.asarray(n_counts, dtype=np.float64)).cuda()
    data = NormedCSR(csr, norm, batch_size=batch_size, feat_idx=feat_idx, n_counts=nc)
    if custom_loadings is not None and data.shape[1] != custom_loadings.shape[0]:
        raise ValueError(f"Provided custom loadings has {custom_loadings.shape[0]} features "
                         f"while the data has {data.shape[1]} features.")
    normed_loc, reduction_loc, ann_loc, knn_loc, graph_loc = graph_locations(
        assay, cell_key, feat_key, dims, pca_cell_key, ann_metric, ann_efc, ann_ef, ann_m,
        rand_state, k, local_connectivity, bandwidth)
    # graph_datastore.py:307-321 and 744: the latest seed partitions' size, else 1000
    if n_centroids is None:
        latest = store.get(reduction_loc + "/latest_kmeans")
        n_centroids = (int(latest.split("/")[-1].split("__")[1]) if latest
                       else DEFAULTS["n_centroids"])
    n_centroids = int(n_centroids)
    kmeans_loc = f"{reduction_loc}/kmeans__{n_centroids}__{rand_state}"
    fit_kmeans = kmeans_loc + "/cluster_centers" not in store
    st = data.stats()
    mu = store.get(normed_loc + "/mu")
    sigma = store.get(normed_loc + "/sigma")
    if mu is None:
        mu = st.mu.cpu().numpy()
        store[normed_loc + "/mu"] = mu
    if sigma is None:
        sigma = st.sigma.cpu().numpy()
        store[normed_loc + "/sigma"] = sigma
    loadings = store.get(reduction_loc + "/reduction")
    if loadings is not None and loadings.shape[0] != data.shape[1]:
        loadings = None
    if custom_loadings is not None:
        loadings = np.asarray(custom_loadings, dtype=np.float64)
    ann_idx = store.get(ann_loc + "/ann_idx")
    ann = AnnStream(data=data, k=k, n_cluster=n_centroids,
                    reduction_method=method, dims=dims, loadings=loadings,
                    use_for_pca=use_for_pca, mu=mu, sigma=sigma, ann_metric=ann_metric,
                    ann_efc=ann_efc, ann_ef=ann_ef, ann_m=ann_m, nthreads=1,
                    ann_parallel=False, rand_state=rand_state, do_kmeans_fit=fit_kmeans,
                    disable_scaling=not feat_scaling, ann_idx=ann_idx, lsi_skip_first=True,
                    lsi_params={}, harmonize=False, screen=screen, pool=pool,
                    pca_mode=pca_mode)
    store[reduction_loc + "/reduction"] = ann.loadings
    store[ann_loc + "/ann_idx"] = ann.annIdx
    if fit_kmeans:
        # graph_datastore.py:958-975: centers and labels, both float64
        store[kmeans_loc + "/cluster_centers"] = ann.kmeans.cluster_centers_.astype(np.float64)
        store[kmeans_loc + "/cluster_labels"] = ann.clusterLabels.astype(np.float64)
    store[reduction_loc + "/latest_kmeans"] = kmeans_loc
    if knn_loc + "/indices" not in store or graph_loc + "/weights" not in store:
        recall = None
        if knn_loc + "/indices" not in store:
            sub = {}
            recall = self_query_knn(ann, sub, batch_size)
            store[knn_loc + "/indices"] = sub["indices"]
            store[knn_loc + "/distances"] = sub["distances"]
            store[knn_loc + "/recall"] = recall
        g = {}
        smoothen_dists(g, store[knn_loc + "/indices"], store[knn_loc + "/distances"],
                       local_connectivity, bandwidth, batch_size)
        store[graph_loc + "/edges"] = g["edges"]
        store[graph_loc + "/weights"] = g["weights"]
    store["latest_graph"] = graph_loc
    store["latest_knn"] = knn_loc
    if return_ann_object:
        return ann
    return store


def load_graph(store: Dict, symmetric: Optional[bool] = None,
               upper_only: Optional[bool] = None, use_k: Optional[int] = None,
               graph_loc: Optional[str] = None):
    """``GraphDataStore.load_graph`` (``graph_datastore.py:1022-1075``): the stored graph as a
    scipy CSR matrix (COO when ``symmetric`` and ``upper_only``), built on the GPU.

    ``graph_loc`` defaults to the latest graph; the cell count and k come from the kNN
    table next to it (``_get_graph_ncells_k``, ``graph_datastore.py:459-472``).
    """
    from .sparse import graph_csr

    if graph_loc is None:
        if "latest_graph" not in store:
            raise ValueError("no graph in the store: run make_graph first")
        graph_loc = store["latest_graph"]
    if graph_loc + "/edges" not in store:
        raise ValueError(f"{graph_loc} not found in the store. Run `make_graph`")
    knn_loc = graph_loc.rsplit("/", 1)[0]
    n_cells, k = np.asarray(store[knn_loc + "/indices"]).shape
    return graph_csr(store[graph_loc + "/edges"], store[graph_loc + "/weights"], n_cells, k,
                     use_k=use_k, symmetric=symmetric is True,
                     upper_only=upper_only is True)


def project_target(ann: AnnStream, target_counts, target_ids: Sequence,
                   source_feat_ids: Sequence, ref_mu: bool = True, ref_sigma: bool = True,
                   target_n_counts=None, run_coral: bool = False):
    """The target cells' embedding (FP32 device [n_target, dims]) as ``run_mapping`` reduces
    them (``mapping_datastore.py:100-208``).

    The target uses the reference's size factor and subset parameters
    (``mapping_datastore.py:106-111``, ``mapping_utils.py:188-194``); features are aligned
    to the reference order with missing ones set to 1.0 (``mapping_utils.py:197-213``).
    ``run_coral`` applies CORAL first (``mapping_datastore.py:160-175``); scarf then maps with
    feature scaling off (``mapping_datastore.py:141-142``), so the corrected data is reduced
    as ``data_coral @ loadings`` with no mu/sigma, whatever ``ref_mu`` / ``ref_sigma`` say.
    """
    require_gpu()
    tcsr = _device_csr(target_counts)
    positions, missing = align_col_map(list(source_feat_ids), list(target_ids))
    if len(positions) != tcsr.n_cols:
        raise ValueError("target_ids must name every column of the target matrix")
    cm = col_map_from_positions(tcsr.n_cols, positions)
    d = len(source_feat_ids)
    norm = ann.data.norm
    libsize = None
    nct = None
    if not norm.renormalize_subset:
        if target_n_counts is None:
            raise ValueError("renormalize_subset=False needs the target's per-cell totals")
        nct = torch.as_tensor(np.asarray(target_n_counts, dtype=np.float64)).cuda()
        libsize = nct
    if run_coral:
        from .coral import coral_projector
        proj, _ = coral_projector(ann, tcsr, cm, missing, norm, n_counts=nct)
        return proj.project(tcsr, norm, libsize=libsize, col_map=cm)
    mu, sigma = ann.mu, ann.sigma
    if not (ref_mu and ref_sigma):
        tst = feature_stats(tcsr, 1000, norm, col_map=cm, d=d, n_counts=nct)
        tm = tst.mu.cpu().numpy()
        ts = tst.sigma.cpu().numpy()
        # aligned columns that are missing in the target are constant 1.0 there
        tm[missing] = 1.0
        ts[missing] = 1.0
        if not ref_mu:
            mu = tm
        if not ref_sigma:
            sigma = ts
    proj = Projector.build(ann.loadings, mu, sigma, clip=norm.clip,
                           disable_scaling=ann._disable_scaling, fill_missing=missing)
    return proj.project(tcsr, norm, libsize=libsize, col_map=cm)


def run_mapping(ann: AnnStream, target_counts, target_ids: Sequence, source_feat_ids: Sequence,
                target_name: str, save_k: int = 3, ref_mu: bool = True, ref_sigma: bool = True,
                target_n_counts=None, store: Optional[Dict] = None, assay: str = "RNA",
                run_coral: bool = False):
    """``MappingDatastore.run_mapping`` (``mapping_datastore.py:31-209``).

    Projects the target (``project_target``), then stores its k = save_k nearest reference
    cells with no self repair (``mapping_datastore.py:198-208``).
    """
    require_gpu()
    store = {} if store is None else store
    if save_k > ann.k:
        save_k = ann.k
    E = project_target(ann, target_counts, target_ids, source_feat_ids, ref_mu=ref_mu,
                       ref_sigma=ref_sigma, target_n_counts=target_n_counts,
                       run_coral=run_coral)
    idx, dist, _ = ann.annIdx.knn_query_tensor(E, save_k)
    loc = f"{assay}/projections/{target_name}"
    store[loc + "/indices"] = idx.cpu().numpy().astype(np.uint64)
    store[loc + "/distances"] = dist.cpu().numpy().astype(np.float64)
    store[loc + "/_ref_n_cells"] = ann.nCells
    return store


def get_mapping_score(store: Dict, target_name: str, target_groups=None,
                      log_transform: bool = True, multiplier: float = 1000,
                      weighted: bool = True, fixed_weight: float = 0.1, assay: str = "RNA",
                      ref_n_cells: Optional[int] = None):
    """Generator of (group, scores) like ``get_mapping_score`` (``mapping_datastore.py:211-285``)."""
    loc = f"{assay}/projections/{target_name}"
    if loc + "/indices" not in store:
        raise KeyError(f"ERROR: Projections have not been computed for {target_name}")
    idx = torch.from_numpy(np.asarray(store[loc + "/indices"]).astype(np.int64)).cuda()
    dist = torch.from_numpy(np.asarray(store[loc + "/distances"], dtype=np.float32)).cuda()
    n_ref = int(ref_n_cells if ref_n_cells is not None else store[loc + "/_ref_n_cells"])
    for g, ms in mapping_scores(idx, dist, n_ref, groups=target_groups,
                                log_transform=log_transform, multiplier=multiplier,
                                weighted=weighted, fixed_weight=fixed_weight):
        yield g, ms


def get_target_classes(store: Dict, target_name: str, reference_class_group,
                       threshold_fraction: float = 0.5, target_subset=None, na_val="NA",
                       assay: str = "RNA"):
    """Class of each target cell by weighted vote of its reference neighbours, like
    ``get_target_classes`` (``mapping_datastore.py:287-366``); ``reference_class_group`` is
    the reference cells' class labels (one per reference cell, as scarf fetches them from
    the cell table).  Returns a pandas Series of labels, ``na_val`` where no class has more
    than ``threshold_fraction`` of the vote."""
    import pandas as pd
    from .mapping import target_classes
    loc = f"{assay}/projections/{target_name}"
    if loc + "/indices" not in store:
        raise KeyError(f"ERROR: Projections have not been computed for {target_name} in th "
                       f"latest graph. Please run `run_mapping` or update latest_graph by "
                       f"running `make_graph` with desired parameters")
    if reference_class_group is None:
        raise ValueError("ERROR: A value is required for the parameter "
                         "`reference_class_group`.")
    if threshold_fraction < 0 or threshold_fraction > 1:
        raise ValueError("ERROR: `threshold_fraction` should have a value between 0 and 1")
    if target_subset is not None and type(target_subset) != list:
        raise TypeError("ERROR:  `target_subset` should be <list> type")
    labels, codes = np.unique(np.asarray(reference_class_group), return_inverse=True)
    idx = torch.from_numpy(np.asarray(store[loc + "/indices"]).astype(np.int64)).cuda()
    dist = torch.from_numpy(np.asarray(store[loc + "/distances"], dtype=np.float64)).cuda()
    res = target_classes(idx, dist, torch.from_numpy(codes.astype(np.int32)),
                         threshold_fraction).cpu().numpy()
    rows = range(res.shape[0])
    if target_subset is not None:
        keep = set(target_subset)
        rows = [n for n in rows if n in keep]
    return pd.Series([labels[res[n]] if res[n] >= 0 else na_val for n in rows])


def load_unified_graph(store: Dict, target_names, use_k: int, target_weight: float,
                       graph_loc: Optional[str] = None, assay: str = "RNA"):
    """Reference kNN graph plus the first ``use_k`` edges of every projected cell of
    ``target_names`` at a uniform ``target_weight``, as a CSR matrix
    (``load_unified_graph``, ``mapping_datastore.py:368-412``).  Returns
    (cell counts [reference, targets...], scipy CSR)."""
    from scipy.sparse import csr_matrix
    if graph_loc is None:
        graph_loc = store["latest_graph"]
    edges = np.asarray(store[graph_loc + "/edges"])
    weights = np.asarray(store[graph_loc + "/weights"])
    ref_n = int(store[f"{assay}/projections/{target_names[0]}/_ref_n_cells"])
    pidx = np.vstack([np.asarray(store[f"{assay}/projections/{x}/indices"])[:, :use_k]
                      for x in target_names])
    n_cells = [ref_n] + [int(np.asarray(store[f"{assay}/projections/{x}/indices"]).shape[0])
                         for x in target_names]
    rows = (ref_n + np.repeat(np.arange(pidx.shape[0]), pidx.shape[1])).astype(np.int64)
    ne = np.stack([rows, pidx.ravel().astype(np.int64)], axis=1)
    me = np.vstack([edges.astype(np.int64), ne])
    mw = np.hstack([weights, np.full(ne.shape[0], float(target_weight))])
    tot = ref_n + pidx.shape[0]
    return n_cells, csr_matrix((mw, (me[:, 0], me[:, 1])), shape=(tot, tot))
