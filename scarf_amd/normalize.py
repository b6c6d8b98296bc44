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


@dataclasses.dataclass
class HostCSR:
    """CSR counts in pinned host memory (torch tensors), ready for asynchronous copies."""

    vals: "torch.Tensor"  # int32 bit patterns of the u32 counts
    cols: "torch.Tensor"  # int32
    rowptr: "torch.Tensor"  # int64, absolute offsets
    n_cols: int

    @property
    def n_rows(self) -> int:
        return int(self.rowptr.shape[0] - 1)

    @classmethod
    def pinned_from(cls, vals, cols, rowptr, n_cols) -> "HostCSR":
        """Copies (host numpy or device tensors) into newly allocated pinned tensors."""
        def pin(x, dt):
            t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))
            out = torch.empty(t.shape, dtype=dt, pin_memory=True)
            out.copy_(t.view(dt) if t.dtype != dt and t.element_size() == out.element_size()
                      else t)
            return out
        return cls(pin(vals, torch.int32), pin(cols, torch.int32), pin(rowptr, torch.int64),
                   int(n_cols))


def upload_with_stats(host: HostCSR, chunk_rows: int, params: NormParams, col_map=None,
                      d: Optional[int] = None, device="cuda", copy_stream=None,
                      stream=None):
    """Streams pinned host CSR counts to the GPU in row chunks and computes the feature
    statistics as the chunks land (the pass scarf makes over its blocks for mean and
    std, ``graph_datastore.py:767-796``): chunk c + 1's copy (on ``copy_stream``)
    overlaps chunk c's statistics kernel (on ``stream``).  Returns the resident
    ``DeviceCSR`` and the same ``FeatureStats`` as :func:`feature_stats` with this
    chunking."""
    require_gpu()
    if not params.renormalize_subset:
        raise NotImplementedError("streamed upload computes the subset library size; pass "
                                  "per-cell totals to feature_stats instead")
    d = host.n_cols if d is None else int(d)
    n = host.n_rows
    nnz = int(host.rowptr[-1])
    vals = torch.empty(max(nnz, 1), dtype=torch.int32, device=device)
    cols = torch.empty(max(nnz, 1), dtype=torch.int32, device=device)
    rowptr = torch.empty(n + 1, dtype=torch.int64, device=device)
    bounds = chunk_bounds(n, chunk_rows)
    nc = len(bounds)
    sums = torch.empty((nc, d), dtype=torch.float64, device=device)
    sqs = torch.empty((nc, d), dtype=torch.float64, device=device)
    counts = torch.tensor([e - s for s, e in bounds], dtype=torch.int64, device=device)
    libsize = torch.empty(max(n, 1), dtype=torch.float64, device=device)
    comp = stream if stream is not None else torch.cuda.current_stream()
    cs = copy_stream if copy_stream is not None else torch.cuda.Stream()
    st = stream_handle(comp)
    rp_h = host.rowptr
    with torch.cuda.stream(comp):
        ws = workspace("stats_stream")
    for c, (a, b) in enumerate(bounds):
        e0, e1 = int(rp_h[a]), int(rp_h[b])
        if c == 0:  # the device arrays were allocated on the compute stream
            cs.wait_stream(comp)
        with torch.cuda.stream(cs):
            vals[e0:e1].copy_(host.vals[e0:e1], non_blocking=True)
            cols[e0:e1].copy_(host.cols[e0:e1], non_blocking=True)
            rowptr[a:b + 1].copy_(rp_h[a:b + 1], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(cs)
        comp.wait_event(ev)
        if b <= a:
            continue
        buf, nbytes = ws.get(_lib.lib().sh_norm_stats_workspace(b - a, b - a, d))
        _lib.call("sh_norm_stats", ptr(vals), ptr(cols), ptr(rowptr, a), b - a, b - a,
                  ptr(col_map), host.n_cols, d, float(params.size_factor),
                  int(params.log_transform), None, ptr(libsize, a), ptr(sums, c * d),
                  ptr(sqs, c * d),
                  ptr(buf), nbytes, st)
    # the device arrays are used on the compute stream from here on
    for t in (vals, cols, rowptr):
        t.record_stream(comp)
    mu = torch.empty(d, dtype=torch.float64, device=device)
    sigma = torch.empty(d, dtype=torch.float64, device=device)
    cmean = torch.empty((nc, d), dtype=torch.float64, device=device)
    _lib.call("sh_stats_merge", ptr(sums), ptr(sqs), ptr(counts), nc, d, ptr(mu), ptr(sigma),
              ptr(cmean), st)
    csr = DeviceCSR(vals[:nnz], cols[:nnz], rowptr, int(host.n_cols))
    return csr, FeatureStats(mu, sigma, libsize[:n], bounds, sums, sqs, cmean, d)
