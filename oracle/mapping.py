"""Restatement of scarf's mapping path (oracle only).

References (scarf source tree):
* ``mapping_utils.py:98-145``   _order_features: for each source feature (in source
                                order) the target column holding it, or -1
* ``mapping_utils.py:148-214``  align_features: the target is normalised over its
                                present features (sorted target order, the source's
                                subset_params) and written in source feature order;
                                missing features are filled with 1.0 (line 208)
* ``mapping_datastore.py:106-111``  the target uses the source's size factor
* ``mapping_datastore.py:176-208``  z with the reference mu/sigma, reducer, then
                                knn_query(k=save_k) without self repair
* ``mapping_datastore.py:211-285``  get_mapping_score
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import norm


def order_features(source_ids, target_ids):
    """Target column for each source feature id (``mapping_utils.py:98-145``, no filters)."""
    t_map = {v: i for i, v in enumerate(target_ids)}
    return np.array([t_map.get(x, -1) for x in source_ids], dtype=np.int64)


def align_normed(target_counts_dense: np.ndarray, t_idx: np.ndarray, sf: float = 1000.0,
                 log_transform: bool = True, renormalize_subset: bool = True,
                 n_counts: Optional[np.ndarray] = None) -> np.ndarray:
    """Aligned FP64 normalised target matrix in source feature order."""
    t_idx = np.asarray(t_idx, dtype=np.int64)
    present = t_idx != -1
    sorted_t = np.array(sorted(t_idx[present]), dtype=np.int64)
    nd = norm.normed(target_counts_dense, sf, log_transform, feat_idx=sorted_t,
                     renormalize_subset=renormalize_subset, n_counts=n_counts)
    unsorter = np.argsort(np.argsort(t_idx[present]))
    out = np.ones((target_counts_dense.shape[0], t_idx.shape[0]))
    out[:, np.where(present)[0]] = nd[:, unsorter]
    return out


def mapping_scores(indices, dists, ref_n_cells: int, target_groups=None,
                   log_transform: bool = True, multiplier: float = 1000,
                   weighted: bool = True, fixed_weight: float = 0.1):
    """``mapping_datastore.py:255-285``: list of (group, scores[ref_n_cells])."""
    indices = np.asarray(indices)
    w = 1 / (np.log1p(np.asarray(dists, dtype=np.float64)) + 1)
    n_cells, n_k = indices.shape
    groups = np.zeros(n_cells) if target_groups is None else np.asarray(target_groups)
    out = []
    for g in sorted(np.unique(groups)):
        sel = np.where(groups == g)[0]
        ms = np.zeros(ref_n_cells)
        for n in sel:
            for x, y in zip(indices[n], w[n]):
                ms[int(x)] += y if weighted else fixed_weight
        ms = multiplier * ms / (len(sel) * n_k)
        if log_transform:
            ms = np.log1p(ms)
        out.append((g, ms))
    return out


def target_classes(indices, dists, ref_groups, threshold_fraction=0.5, target_subset=None,
                   na_val="NA"):
    """Label transfer, restating ``get_target_classes`` (``mapping_datastore.py:340-366``)
    line for line: weights ``1 - d / max(d)`` per row, the first k - 1 neighbours vote in
    neighbour order, the unique class above the threshold wins, else ``na_val``."""
    indices = np.asarray(indices)
    dists = np.asarray(dists, dtype=np.float64)
    if target_subset is not None:
        target_subset = {x: None for x in target_subset}
    preds = []
    with np.errstate(divide="ignore", invalid="ignore"):
        weights = 1 - (dists / dists.max(axis=1).reshape(-1, 1))
    for n in range(indices.shape[0]):
        if target_subset is not None and n not in target_subset:
            continue
        wd = {}
        for i, j in zip(indices[n, :-1], weights[n, :-1]):
            k = ref_groups[i]
            if k not in wd:
                wd[k] = 0
            wd[k] += j
        temp = na_val
        s = weights[n, :-1].sum()
        for i, j in wd.items():
            if j / s > threshold_fraction:
                if temp == na_val:
                    temp = i
                else:
                    temp = na_val
                    break
        preds.append(temp)
    return preds
