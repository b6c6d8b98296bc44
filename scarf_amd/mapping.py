"""Mapping of query cells onto a reference (scarf ``run_mapping`` / ``get_mapping_score``).

``align_col_map`` reproduces ``align_features`` (``mapping_utils.py:98-214``) as a
column map for the GPU kernels: each target column that holds a source feature is
mapped to that feature's position; the target's library size is taken over those
columns (the source's ``subset_params``), and source features missing from the target
are filled with the normalised value 1.0 (``mapping_utils.py:208``), which the
projection folds into its constant row.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _lib
from .device import ptr, require_gpu, stream_handle, workspace

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def align_col_map(source_feat_ids, target_ids):
    """(positions int32 [n_target_cols], missing bool [d_source]) for the target matrix."""
    s_pos = {v: i for i, v in enumerate(source_feat_ids)}
    positions = np.array([s_pos.get(t, -1) for t in target_ids], dtype=np.int32)
    present = np.zeros(len(source_feat_ids), dtype=bool)
    present[positions[positions >= 0]] = True
    if present.sum() == 0:
        raise ValueError("ERROR: None of the features from reference were found in the target data")
    return positions, ~present


def mapping_scores(idx: "torch.Tensor", dist: "torch.Tensor", n_ref: int, groups=None,
                   log_transform: bool = True, multiplier: float = 1000.0,
                   weighted: bool = True, fixed_weight: float = 0.1, stream=None):
    """List of (group, scores f64 [n_ref]) like ``get_mapping_score``'s generator."""
    require_gpu()
    nq, k = idx.shape
    dev = idx.device
    if groups is not None:
        groups = np.asarray(groups)
        if groups.shape[0] != nq:
            raise ValueError(f"ERROR: Length of target_groups {groups.shape[0]} not same as "
                             f"number of target cells in the projection {nq}")
        labels, codes = np.unique(groups, return_inverse=True)
        g_t = torch.from_numpy(codes.astype(np.int32)).to(dev)
    else:
        labels, g_t = np.array([0.0]), None
    ng = len(labels)
    out = torch.empty((ng, n_ref), dtype=torch.float64, device=dev)
    buf, nbytes = workspace("mapping").get(_lib.lib().sh_mapping_score_workspace(n_ref, ng))
    idx, dist = idx.contiguous(), dist.contiguous()  # referenced until the launch is queued
    _lib.call("sh_mapping_score", ptr(idx), ptr(dist), nq, k,
              ptr(g_t), ng, n_ref, int(weighted), float(fixed_weight), float(multiplier),
              int(log_transform), ptr(out), ptr(buf), nbytes, stream_handle(stream))
    o = out.cpu().numpy()
    return [(labels[i], o[i]) for i in range(ng)]


def target_classes(idx: "torch.Tensor", dist: "torch.Tensor", ref_codes: "torch.Tensor",
                   threshold: float = 0.5, stream=None) -> "torch.Tensor":
    """Label transfer on the GPU (``sh_target_classes``): int32 class code per target row,
    -1 where no single reference class holds more than ``threshold`` of the vote
    (``get_target_classes``, ``mapping_datastore.py:287-366``).  ``dist`` is FP64, as the
    projections store holds it; ``ref_codes`` int32 per reference cell."""
    require_gpu()
    n, k = idx.shape
    out = torch.empty(max(n, 1), dtype=torch.int32, device=idx.device)
    idx = idx.to(torch.int64).contiguous()
    dist = dist.to(torch.float64).contiguous()
    ref_codes = ref_codes.to(device=idx.device, dtype=torch.int32).contiguous()
    _lib.call("sh_target_classes", ptr(idx), ptr(dist), n, k, ptr(ref_codes),
              int(ref_codes.numel()), float(threshold), ptr(out), stream_handle(stream))
    return out[:n]
