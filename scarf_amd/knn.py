"""Exact kNN on the GPU behind an hnswlib-compatible index object.

scarf builds an ``hnswlib.Index(space='l2')`` (``ann.py:14-28``), fills it block by
block with ``add_items`` (``ann.py:322-325``) and queries it with ``knn_query``
(``ann.py:194-205``), which returns uint64 labels and float32 *squared* Euclidean
distances in ascending order.  ``GpuIndex`` keeps the reference embedding resident in
HBM and answers ``knn_query`` with an exact search:

* ``screen="fp32"``: FP32 MFMA screen, exact FP64 re-rank of the pool, and an
  error-bound certificate; rows it cannot certify are recomputed by FP64 brute
  force.  The result equals the FP64 brute-force top-k (ties to the lower index).
* ``screen="bf16"``: BF16 MFMA screen (FP32 accumulation) keeping a ``pool`` of
  candidates per query, re-ranked exactly; recall against FP64 is measured.

Self-queries (``self_offset``) return k neighbours after scarf's ``fix_knn_query``
repair (``ann.py:31-52``).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

from . import _lib
from .device import ptr, require_gpu, stream_handle, workspace

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

U32 = 2.0 ** -24

# Optional per-launch HIP-event timing of the hot kernels (enabled by measurement
# harnesses): name -> list of (start event, end event, algorithmic work).
_timing = None


def enable_kernel_timing():
    global _timing
    _timing = {}
    return _timing


def disable_kernel_timing():
    global _timing
    t, _timing = _timing, None
    return t


class kernel_timer:
    def __init__(self, name, stream=None, flops=0.0, nbytes=0.0):
        self.name, self.stream, self.flops, self.nbytes = name, stream, flops, nbytes

    def __enter__(self):
        if _timing is not None:
            self.s = torch.cuda.Event(enable_timing=True)
            self.s.record(self.stream or torch.cuda.current_stream())
        return self

    def __exit__(self, *exc):
        if _timing is not None:
            e = torch.cuda.Event(enable_timing=True)
            e.record(self.stream or torch.cuda.current_stream())
            _timing.setdefault(self.name, []).append((self.s, e, self.flops, self.nbytes))
        return False


@dataclasses.dataclass
class PackedRefs:
    dtype: str
    D: int
    K: int
    data: "torch.Tensor"  # packed operand rows
    norms: "torch.Tensor"  # f32 ||r||^2
    norm_max: float


def pack(E: "torch.Tensor", dtype: str, role: int, stream=None):
    """Pack an FP32 embedding [n, D] into the screen operand layout."""
    require_gpu()
    code = _lib.DTYPES[dtype]
    n, D = E.shape
    K = int(_lib.lib().sh_knn_pack_k(code, D))
    rowb = int(_lib.lib().sh_knn_pack_row_bytes(code, D))
    if K <= 0 or rowb <= 0:
        raise ValueError(f"screen dtype {dtype} unsupported for D={D}")
    out = torch.empty((max(n, 1), rowb), dtype=torch.uint8, device=E.device)
    aux = torch.empty(max(n, 1), dtype=torch.float32, device=E.device)
    _lib.call("sh_knn_pack", ptr(E), n, D, E.stride(0), code, role, ptr(out), ptr(aux),
              stream_handle(stream))
    return out[:n], aux[:n], K


def pack_refs(E: "torch.Tensor", dtype: str, stream=None) -> PackedRefs:
    data, norms, K = pack(E, dtype, 0, stream)
    nm = float(torch.sqrt(norms.max()).item()) if norms.numel() else 0.0
    return PackedRefs(dtype, int(E.shape[1]), K, data, norms, nm)


def default_slices(nq: int, nr: int) -> int:
    """Reference slices so that small query batches still fill the GPU (>= ~512 WGs)."""
    qblocks = max(1, math.ceil(nq / 256))
    want = max(1, math.ceil(512 / qblocks))
    max_by_len = max(1, nr // 2048)  # keep slices >= 2048 references
    return int(min(want, max_by_len, 4096))


def _fix_slices(nr: int, s: int) -> int:
    # the kernel requires every slice to hold >= 1 reference after 128-alignment
    while s > 1:
        sl = int(math.ceil(math.ceil(nr / s) / 128) * 128)
        if math.ceil(nr / sl) == s:
            return s
        s -= 1
    return 1


# query batches at least this large are screened in locality order
ORDER_MIN_QUERIES = 8192


def locality_order(E: "torch.Tensor", stream=None) -> "torch.Tensor":
    """Rows of ``E`` (FP32 [n, D] on device) in Morton order of their first three
    coordinates (sh_locality_order), as an int64 permutation."""
    n, D = E.shape
    perm = torch.empty(n, dtype=torch.int64, device=E.device)
    if n == 0:
        return perm
    buf, nbytes = workspace("order").get(_lib.lib().sh_locality_order_workspace(n))
    _lib.call("sh_locality_order", ptr(E), n, D, E.stride(0), ptr(perm), ptr(buf), nbytes,
              stream_handle(stream))
    return perm


def permute_rows(src: "torch.Tensor", perm: "torch.Tensor", scatter: bool = False,
                 out: Optional["torch.Tensor"] = None, stream=None) -> "torch.Tensor":
    """Row gather (``out[i] = src[perm[i]]``) or scatter (``out[perm[i]] = src[i]``) of a
    2-D device tensor whose rows are a whole number of 4-byte words."""
    n = int(perm.numel())
    if src.dim() == 1:
        src = src.view(-1, 1)
    if out is None:
        out = torch.empty((n if not scatter else src.shape[0],) + tuple(src.shape[1:]),
                          dtype=src.dtype, device=src.device)
    o2 = out.view(out.shape[0], -1)
    es = src.element_size()
    _lib.call("sh_permute_rows", ptr(src), src.stride(0) * es, ptr(perm), n,
              src.shape[1] * es, ptr(o2), o2.stride(0) * es, int(scatter), stream_handle(stream))
    return out


# Margin test of the FP8 screen (sh_knn_rerank's SH_FLAG_THIN) and the pool of the second
# screen that the flagged queries get.  On the 4M-cell, 100-dimension embedding, margin
# 1.0 sends 8.7% of the queries to the second screen and lifts recall@11 from 0.9977 to
# 0.9993 for 20% more kNN time (0.75: 5.2%, 0.9989; 1.5: 17.5%, 0.9994).
FP8_MARGIN = 1.0
REFINE_POOL = 256


def search(Q: "torch.Tensor", R: "torch.Tensor", refs: PackedRefs, k: int, pool: int = 64,
           self_offset: Optional[int] = None, n_slices: Optional[int] = None,
           certify: Optional[bool] = None, self_idx: Optional["torch.Tensor"] = None,
           order: Optional[bool] = None, seed: bool = True, refine: Optional[bool] = None,
           stream=None):
    """Top-k of each query row of ``Q`` against ``R`` (both FP32 [*, D] on device).

    Returns (idx int64 [nq, k], dist f32 [nq, k], flags int32 [nq]).  ``order`` (default:
    batches of at least ORDER_MIN_QUERIES rows) screens the queries in locality order;
    with ``self_offset`` and ``seed``, each block of locality-ordered queries also seeds
    the screen threshold with its own reference rows.  The results are the same either
    way.

    ``refine`` (default: on for the FP8 screen below a pool of REFINE_POOL): the re-rank
    flags queries whose k-th exact distance is not clear of the pool's cut-off by
    FP8_MARGIN times the largest screening error seen in the pool (SH_FLAG_THIN); those
    queries are screened again with a pool of REFINE_POOL and their rows replaced.
    """
    require_gpu()
    if refine is None:
        refine = refs.dtype == "fp8" and pool < REFINE_POOL
    margin = FP8_MARGIN if refine else 0.0
    idx, dist, flags = _search_ordered(Q, R, refs, k, pool, self_offset, n_slices, certify,
                                       self_idx, order, seed, margin, stream)
    if refine:
        thin = torch.nonzero(flags & _lib.SH_FLAG_THIN).flatten()
        if thin.numel():
            if self_idx is not None:
                si = self_idx.to(device=Q.device, dtype=torch.int64)[thin]
            elif self_offset is not None:
                si = thin + int(self_offset)
            else:
                si = None
            i2, d2, f2 = _search_ordered(Q[thin], R, refs, k, REFINE_POOL, None, n_slices,
                                         certify, si, None, seed, 0.0, stream)
            idx[thin] = i2
            dist[thin] = d2
            flags[thin] = f2 | _lib.SH_FLAG_THIN
    return idx, dist, flags


def _search_ordered(Q, R, refs, k, pool, self_offset, n_slices, certify, self_idx, order, seed,
                    margin, stream):
    nq = Q.shape[0]
    if order is None:
        order = nq >= ORDER_MIN_QUERIES
    if not order or nq == 0:
        return _search(Q, R, refs, k, pool, self_offset, n_slices, certify, self_idx, stream,
                       margin=margin)
    perm = locality_order(Q, stream)
    Qs = permute_rows(Q, perm, stream=stream)
    seeds = None
    if self_idx is not None:
        si = self_idx.to(device=Q.device, dtype=torch.int64).contiguous()
        if si.numel() != nq:
            raise ValueError("self_idx needs one reference index per query row")
        si_s = permute_rows(si, perm, stream=stream).view(-1)
    elif self_offset is not None:
        si_s = perm + int(self_offset)
        if seed:  # distinct rows of R, near each other in locality order
            seeds = si_s
    else:
        si_s = None
    idx_s, dist_s, flags_s = _search(Qs, R, refs, k, pool, None, n_slices, certify, si_s,
                                     stream, seeds=seeds, margin=margin)
    idx = permute_rows(idx_s, perm, scatter=True, stream=stream)
    dist = permute_rows(dist_s, perm, scatter=True, stream=stream)
    flags = permute_rows(flags_s.view(-1, 1), perm, scatter=True, stream=stream).view(-1)
    return idx, dist, flags


def _search(Q, R, refs, k, pool, self_offset, n_slices, certify, self_idx, stream, seeds=None,
            margin=0.0):
    nq, D = Q.shape
    nr = R.shape[0]
    dev = Q.device
    if D != refs.D:
        raise ValueError(f"query dim {D} != index dim {refs.D}")
    has_self = self_offset is not None or self_idx is not None
    if self_idx is not None:
        self_idx = self_idx.to(device=Q.device, dtype=torch.int64).contiguous()
        if self_idx.numel() != nq:
            raise ValueError("self_idx needs one reference index per query row")
    kk = k + (1 if has_self else 0)
    if kk > nr:
        raise ValueError(f"k={k} needs at least {kk} reference rows, index has {nr}")
    pool = max(int(pool), kk)
    if pool > 256 or (pool > 128 and refs.dtype != "fp8"):
        raise ValueError("pool above 128 is supported for the FP8 screen only, up to 256")
    idx = torch.empty((nq, k), dtype=torch.int64, device=dev)
    dist = torch.empty((nq, k), dtype=torch.float32, device=dev)
    flags = torch.zeros(max(nq, 1), dtype=torch.int32, device=dev)
    if nq == 0:
        return idx, dist, flags[:0]
    code = _lib.DTYPES[refs.dtype]
    st = stream_handle(stream)
    qp, _, _ = pack(Q, refs.dtype, 1, stream)
    S = _fix_slices(nr, n_slices if n_slices else default_slices(nq, nr))
    if S > 1:  # seeds pay off when every block sees all of its seed rows
        seeds = None
    cand = torch.empty((nq, S, pool), dtype=torch.int64, device=dev)
    cnt = torch.empty((nq, S), dtype=torch.int32, device=dev)
    ws, nbytes = workspace("screen").get(_lib.lib().sh_knn_screen_workspace(nq, pool, S))
    with kernel_timer("screen", stream, flops=2.0 * nq * nr * D):
        _lib.call("sh_knn_screen", ptr(qp), nq, ptr(refs.data), nr, D, code, None, None, pool,
                  S, ptr(seeds), ptr(cand), ptr(cnt), ptr(ws), nbytes, st)
    if certify is None:
        certify = refs.dtype in ("fp32", "f32")
    err_coef = -1.0
    if certify:
        # |s~ - s| <= (K + 2) u (||r||^2 + 2 |q| |r|) for the FP32 screen; x2 margin
        err_coef = 2.0 * (refs.K + 2) * U32
    so = -1 if self_offset is None else int(self_offset)
    _lib.call("sh_knn_rerank", ptr(Q), nq, Q.stride(0), ptr(R), nr, R.stride(0), D, ptr(cand),
              ptr(cnt), pool, S, k, so, ptr(self_idx), float(err_coef), float(refs.norm_max),
              float(margin), ptr(idx), ptr(dist), ptr(flags), st)
    if certify:
        bad = torch.nonzero(flags[:nq] & _lib.SH_FLAG_UNCERTIFIED).flatten()
        if bad.numel():
            exact_rows(Q, R, k, bad, idx, dist, flags, self_offset, stream, self_idx=self_idx)
    return idx, dist, flags[:nq]


def exact_rows(Q, R, k, rows, idx, dist, flags, self_offset=None, stream=None, dist64=None,
               self_idx=None):
    """FP64 brute force for the query rows listed in ``rows`` (device int64)."""
    nq, D = Q.shape
    so = -1 if self_offset is None else int(self_offset)
    _lib.call("sh_knn_exact", ptr(Q), nq, Q.stride(0), ptr(rows), int(rows.numel()), ptr(R),
              R.shape[0], R.stride(0), D, k, so, ptr(self_idx), ptr(idx), ptr(dist), ptr(dist64),
              ptr(flags), stream_handle(stream))


def fix_knn_query_device(idx, dist, ref_idx, stream=None):
    """fix_knn_query (ann.py:31-52) on device tables: (idx [n, k], dist [n, k], n_mis)."""
    require_gpu()
    n, k1 = idx.shape
    dev = idx.device
    oi = torch.empty((n, k1 - 1), dtype=torch.int64, device=dev)
    od = torch.empty((n, k1 - 1), dtype=torch.float32, device=dev)
    fl = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
    # contiguous copies stay referenced until the launch is queued
    idx, dist, ref_idx = idx.contiguous(), dist.contiguous(), ref_idx.contiguous()
    _lib.call("sh_fix_knn_query", ptr(idx), ptr(dist), ptr(ref_idx), n, k1, ptr(oi), ptr(od),
              ptr(fl), stream_handle(stream))
    n_mis = int(((fl[:n] & _lib.SH_FLAG_SELF_NOT_FIRST) > 0).sum().item())
    return oi, od, n_mis


def exact_knn(Q, R, k, self_offset=None, stream=None):
    """FP64 brute force for every query row (the GPU reference for recall)."""
    require_gpu()
    nq = Q.shape[0]
    dev = Q.device
    idx = torch.empty((nq, k), dtype=torch.int64, device=dev)
    dist = torch.empty((nq, k), dtype=torch.float32, device=dev)
    d64 = torch.empty((nq, k), dtype=torch.float64, device=dev)
    flags = torch.zeros(max(nq, 1), dtype=torch.int32, device=dev)
    so = -1 if self_offset is None else int(self_offset)
    _lib.call("sh_knn_exact", ptr(Q), nq, Q.stride(0), None, nq, ptr(R), R.shape[0],
              R.stride(0), Q.shape[1], k, so, None, ptr(idx), ptr(dist), ptr(d64), ptr(flags),
              stream_handle(stream))
    return idx, dist, d64, flags[:nq]


class GpuIndex:
    """hnswlib ``Index`` duck type (space 'l2') with an exact GPU search.

    Constructor ``(space, dim)``; ``init_index``, ``add_items``, ``knn_query``,
    ``set_ef``, ``set_num_threads``, ``save_index``, ``load_index``,
    ``get_current_count`` behave as scarf uses them (``ann.py:14-28``,
    ``graph_datastore.py:875-881, 945-947``).  ``ef``/``M``/threads are accepted and
    recorded; they do not change the (exact) result.
    """

    def __init__(self, space: str = "l2", dim: int = 0, screen: str = "bf16", pool: int = 64,
                 device: str = "cuda"):
        if space != "l2":
            raise ValueError(f"only space='l2' is supported (got {space!r})")
        self.space = space
        self.dim = int(dim)
        self.screen = screen
        self.pool = int(pool)
        self.device = device
        self.max_elements = 0
        self.ef = 10
        self.ef_construction = 200
        self.M = 16
        self.random_seed = 100
        self.num_threads = 1
        self._chunks = []
        self._E = None
        self._packed = None

    # -- construction -----------------------------------------------------------
    def init_index(self, max_elements: int, ef_construction: int = 200, M: int = 16,
                   random_seed: int = 100, allow_replace_deleted: bool = False):
        self.max_elements = int(max_elements)
        self.ef_construction = int(ef_construction)
        self.M = int(M)
        self.random_seed = int(random_seed)

    def set_ef(self, ef: int):
        self.ef = int(ef)

    def set_num_threads(self, n: int):
        self.num_threads = int(n)

    def _as_f32(self, data):
        if torch is not None and isinstance(data, torch.Tensor):
            t = data.to(device=self.device, dtype=torch.float32)
        else:
            a = np.ascontiguousarray(np.asarray(data), dtype=np.float32)
            if a.ndim == 1:
                a = a[None, :]
            t = torch.from_numpy(a).to(self.device)
        if t.ndim == 1:
            t = t[None, :]
        if t.shape[1] != self.dim:
            raise RuntimeError(f"Wrong dimensionality of the vectors: {t.shape[1]} != {self.dim}")
        return t.contiguous()

    def add_items(self, data, ids=None, num_threads: int = -1, replace_deleted=False):
        require_gpu()
        if ids is not None:
            ids = np.asarray(ids).ravel()
            start = self.get_current_count()
            if not np.array_equal(ids, np.arange(start, start + ids.shape[0])):
                raise ValueError("GpuIndex labels are the insertion order; custom ids unsupported")
        t = self._as_f32(data)
        if self.max_elements and self.get_current_count() + t.shape[0] > self.max_elements:
            raise RuntimeError("The number of elements exceeds the specified limit")
        self._chunks.append(t)
        self._E = None
        self._packed = None

    def add_items_tensor(self, E: "torch.Tensor"):
        """Adds rows already resident on the device without a host round trip."""
        self.add_items(E)

    def get_current_count(self) -> int:
        if self._E is not None:
            return int(self._E.shape[0])
        return int(sum(c.shape[0] for c in self._chunks))

    @property
    def element_count(self) -> int:
        return self.get_current_count()

    def embedding(self) -> "torch.Tensor":
        if self._E is None:
            if not self._chunks:
                return torch.empty((0, self.dim), dtype=torch.float32, device=self.device)
            self._E = torch.cat(self._chunks) if len(self._chunks) > 1 else self._chunks[0]
            self._chunks = [self._E]
        return self._E

    def packed(self) -> PackedRefs:
        if self._packed is None:
            self._packed = pack_refs(self.embedding(), self.screen)
        return self._packed

    # -- query --------------------------------------------------------------------
    def knn_query_tensor(self, Q: "torch.Tensor", k: int = 1, self_offset=None,
                         self_idx=None):
        """Device in, device out: (int64 idx, f32 squared distances, flags)."""
        return search(Q, self.embedding(), self.packed(), k, self.pool, self_offset,
                      self_idx=self_idx)

    def knn_query(self, data, k: int = 1, num_threads: int = -1, filter=None):
        require_gpu()
        n = self.get_current_count()
        if k > n:
            raise RuntimeError("Cannot return the results in a contiguous 2D array. "
                               "Probably ef or M is too small")
        Q = self._as_f32(data)
        idx, dist, _ = self.knn_query_tensor(Q, k)
        return idx.cpu().numpy().astype(np.uint64), dist.cpu().numpy()

    # -- persistence ----------------------------------------------------------------
    def save_index(self, path: str):
        E = self.embedding().cpu().numpy()
        meta = np.array([self.dim, self.max_elements, self.ef, self.ef_construction, self.M,
                         self.random_seed, self.pool], dtype=np.int64)
        with open(path, "wb") as f:
            np.savez(f, embedding=E, meta=meta, screen=np.array(self.screen))

    def load_index(self, path: str, max_elements: int = 0, allow_replace_deleted=False):
        with np.load(path, allow_pickle=False) as z:
            E = z["embedding"]
            meta = z["meta"]
            self.screen = str(z["screen"])
        self.dim = int(meta[0])
        self.max_elements = max(int(meta[1]), int(max_elements))
        self.ef, self.ef_construction, self.M, self.random_seed, self.pool = \
            (int(x) for x in meta[2:7])
        self._chunks = [torch.from_numpy(np.ascontiguousarray(E, dtype=np.float32)).to(self.device)]
        self._E = None
        self._packed = None
