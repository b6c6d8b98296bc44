"""Library-size normalisation and per-feature statistics on the GPU.

Mirrors ``RNAassay.normed`` (scarf ``assay.py:780-828``) with ``renormalize_subset``
and ``log_transform``, and the mean / std of the normalised data that ``make_graph``
computes and cleans (``graph_datastore.py:767-796``, ``utils.py:143-154``).  The
normalised matrix is never materialised: every later kernel recomputes
x = log1p(sf * c / libsize) from the resident CSR counts.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from . import _lib
from .device import DeviceCSR, ptr, require_gpu, stream_handle, workspace

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


@dataclasses.dataclass
class NormParams:
    """Normalisation semantics (scarf defaults: sf 1000, log1p, subset library size)."""

    size_factor: float = 1000.0
    log_transform: bool = True
    renormalize_subset: bool = True
    clip: Optional[float] = None  # z-score clip; scarf has none (ann.py:191-192)

    def key(self):
        return {"log_transform": self.log_transform,
                "renormalize_subset": self.renormalize_subset}


@dataclasses.dataclass
class FeatureStats:
    mu: "torch.Tensor"  # f64 [d]
    sigma: "torch.Tensor"  # f64 [d]
    libsize: "torch.Tensor"  # f64 [n]
    chunk_bounds: list  # [(start, end)] row ranges
    chunk_sums: "torch.Tensor"  # f64 [n_chunks, d]
    chunk_sumsq: "torch.Tensor"  # f64 [n_chunks, d]
    chunk_mean: "torch.Tensor"  # f64 [n_chunks, d]
    d: int


def chunk_bounds(n_rows: int, chunk_rows: int):
    chunk_rows = max(1, int(chunk_rows))
    return [(s, min(s + chunk_rows, n_rows)) for s in range(0, n_rows, chunk_rows)] or [(0, 0)]


def feature_stats(csr: DeviceCSR, chunk_rows: int, params: NormParams, col_map=None,
                  d: Optional[int] = None, n_counts: Optional["torch.Tensor"] = None,
                  stream=None) -> FeatureStats:
    """Per-chunk FP64 sums on the GPU, merged (Chan) into mu/sigma with clean_array.

    ``n_counts`` (f64 per cell) replaces the subset library size when
    ``params.renormalize_subset`` is False (``assay.py:824-825``).
    """
    require_gpu()
    d = csr.n_cols if d is None else int(d)
    n = csr.n_rows
    dev = csr.rowptr.device
    bounds = chunk_bounds(n, chunk_rows)
    nc = len(bounds)
    sums = torch.empty((nc, d), dtype=torch.float64, device=dev)
    sqs = torch.empty((nc, d), dtype=torch.float64, device=dev)
    counts = torch.tensor([e - s for s, e in bounds], dtype=torch.int64, device=dev)
    libsize = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
    if not params.renormalize_subset:
        if n_counts is None:
            raise ValueError("renormalize_subset=False needs the per-cell total counts")
        libsize[:n] = n_counts.to(device=dev, dtype=torch.float64)
    st = stream_handle(stream)
    buf, nbytes = workspace("stats").get(
        _lib.lib().sh_norm_stats_workspace(n, int(chunk_rows), d))
    _lib.call("sh_norm_stats", ptr(csr.vals), ptr(csr.cols), ptr(csr.rowptr), n, int(chunk_rows),
              ptr(col_map), csr.n_cols, d, float(params.size_factor), int(params.log_transform),
              ptr(libsize) if not params.renormalize_subset else None, ptr(libsize), ptr(sums),
              ptr(sqs), ptr(buf), nbytes, st)
    mu = torch.empty(d, dtype=torch.float64, device=dev)
    sigma = torch.empty(d, dtype=torch.float64, device=dev)
    cmean = torch.empty((nc, d), dtype=torch.float64, device=dev)
    _lib.call("sh_stats_merge", ptr(sums), ptr(sqs), ptr(counts), nc, d, ptr(mu), ptr(sigma),
              ptr(cmean), st)
    return FeatureStats(mu, sigma, libsize[:n], bounds, sums, sqs, cmean, d)
