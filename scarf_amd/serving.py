"""Continuous-batching mapping server over a reference resident in HBM.

Each request is a batch of query cells (CSR counts in the request's own gene order).
The server answers with what scarf's ``run_mapping`` + ``get_mapping_score`` compute
(``mapping_datastore.py:31-285``): for each query cell the k nearest reference cells
and their squared distances, and the request's mapping score of every reference cell.

Scheduling
----------
* Admission by memory: a request is admitted when its device footprint (counts,
  neighbour tables, score vector) fits in what is left of ``mem_budget``; requests
  that could never fit are rejected with an error.  Admission is FIFO.
* Requests are split into chunks of at most ``chunk_rows`` cells.  The engine issues
  chunks round-robin over the admitted requests, each on one of ``n_streams`` HIP
  streams, so a small request that arrives while a large one is on the GPU is admitted
  at once and its chunks interleave with the large one's; chunks on different streams
  overlap on the device.  A request's slot and memory are released when it completes.
* Determinism: every kernel's result for a query row depends only on that row and the
  resident reference (projection per row, canonical kNN pool, 2^-40 fixed-point score
  sums), so a request's answer is bit-identical alone and in any mix.
* Uploads go through per-stream pinned staging buffers (reused; a stream is handed out
  again only after its work has completed), so the host-to-device copies are
  asynchronous and overlap the other streams' kernels.
* Cancellation: ``future.cancel()`` on a request's future.  A queued request is dropped;
  an admitted one issues no further chunks and releases its memory and slot as soon as
  its chunks in flight have completed.
* Idle waiting: when nothing can be issued the engine sleeps on a condition variable
  that a submission or a cancellation wakes, checking its in-flight events every
  ``poll_s`` seconds.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import dataclasses
import threading
import time
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .device import DeviceCSR, col_map_from_positions, ptr, require_gpu, stream_handle, workspace
from .knn import GpuIndex, search
from .mapping import align_col_map
from .normalize import NormParams
from .pca import Projector

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


@dataclasses.dataclass
class MapResult:
    rid: int
    n_cells: int
    indices: Optional[np.ndarray] = None  # uint64 [n, k]
    distances: Optional[np.ndarray] = None  # float32 [n, k] squared L2
    scores: Optional[np.ndarray] = None  # float64 [n_ref]
    arrival: float = 0.0
    admitted: float = 0.0
    done: float = 0.0
    error: Optional[str] = None

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3


class _Req:
    def __init__(self, rid, csr, arrival, fut):
        self.rid = rid
        self.csr = csr
        self.arrival = arrival
        self.fut = fut
        self.n = csr.n_rows
        self.next_row = 0
        self.outstanding = 0
        self.dev = None
        self.idx = None
        self.dist = None
        self.need = 0
        self.admitted = 0.0
        self.res = MapResult(rid, self.n, arrival=arrival)


class MappingServer:
    """Maps query batches onto a resident reference with continuous batching."""

    def __init__(self, loadings: np.ndarray, mu: np.ndarray, sigma: np.ndarray,
                 index: GpuIndex, source_feat_ids: Sequence, query_ids: Sequence,
                 norm: NormParams = NormParams(), k: int = 11, mem_budget: int = 4 << 30,
                 chunk_rows: int = 16384, n_streams: int = 4, multiplier: float = 1000.0,
                 log_transform_scores: bool = True, disable_scaling: bool = False,
                 poll_s: float = 0.0005):
        require_gpu()
        self.index = index
        self.k = int(k)
        self.norm = norm
        self.n_ref = index.get_current_count()
        positions, missing = align_col_map(list(source_feat_ids), list(query_ids))
        self.n_cols = len(query_ids)
        self.col_map = col_map_from_positions(self.n_cols, positions)
        self.proj = Projector.build(loadings, mu, sigma, clip=norm.clip,
                                    disable_scaling=disable_scaling, fill_missing=missing)
        self.refs = index.packed()  # resident packed reference operand
        self.E = index.embedding()
        self.mem_budget = int(mem_budget)
        self.chunk_rows = int(chunk_rows)
        self.multiplier = float(multiplier)
        self.log_scores = bool(log_transform_scores)
        self.streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
        self._staging = [None] * len(self.streams)  # pinned upload buffer per stream
        self._free_streams = collections.deque(range(len(self.streams)))
        self.poll_s = float(poll_s)
        self._queue = collections.deque()
        self._active: List[_Req] = []
        self._inflight = []  # (event, req, kind, stream_id)
        self._lock = threading.Condition()
        self._used = 0
        self._rid = 0
        self._stop = False
        self._thread = None
        self._rr = 0

    # -- client API ---------------------------------------------------------------------
    def footprint(self, csr) -> int:
        """Device bytes a request holds while admitted."""
        n = csr.n_rows
        return (csr.nnz * 8 + (n + 1) * 8 + n * self.k * 12
                + self.n_ref * 16 + 4096)

    def submit(self, csr, arrival: Optional[float] = None) -> cf.Future:
        if csr.n_cols != self.n_cols:
            raise ValueError(f"request has {csr.n_cols} columns, server expects {self.n_cols}")
        fut = cf.Future()
        with self._lock:
            rid = self._rid
            self._rid += 1
            req = _Req(rid, csr, time.perf_counter() if arrival is None else arrival, fut)
            req.need = self.footprint(csr)
            if req.need > self.mem_budget:
                req.res.error = (f"request needs {req.need} device bytes, more than the "
                                 f"admission budget {self.mem_budget}")
                req.res.done = time.perf_counter()
                fut.set_result(req.res)
                return fut
            if csr.n_rows == 0:
                req.res.indices = np.zeros((0, self.k), np.uint64)
                req.res.distances = np.zeros((0, self.k), np.float32)
                req.res.scores = np.zeros(self.n_ref)
                req.res.done = time.perf_counter()
                fut.set_result(req.res)
                return fut
            self._queue.append(req)
            self._lock.notify()
        return fut

    def cancel(self, fut: cf.Future) -> bool:
        """Cancels a submitted request (the same as ``fut.cancel()``)."""
        ok = fut.cancel()
        with self._lock:
            self._lock.notify()
        return ok

    def start(self):
        self._stop = False
        self._thread = threading.Thread(target=self._engine, name="scarf-amd-mapping",
                                        daemon=True)
        self._thread.start()
        return self

    def stop(self):
        with self._lock:
            self._stop = True
            self._lock.notify()
        if self._thread is not None:
            self._thread.join()
        torch.cuda.synchronize()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    def map_one(self, csr) -> MapResult:
        """Synchronous convenience: one request through the running engine."""
        return self.submit(csr).result()

    # -- engine ---------------------------------------------------------------------------
    def _admit(self):
        now = time.perf_counter()
        if any(r.fut.cancelled() for r in self._queue):
            self._queue = collections.deque(r for r in self._queue if not r.fut.cancelled())
        while self._queue and self._used + self._queue[0].need <= self.mem_budget:
            req = self._queue.popleft()
            self._used += req.need
            req.admitted = now
            req.res.admitted = now
            self._active.append(req)

    def _pick(self) -> Optional[_Req]:
        n = len(self._active)
        for off in range(n):
            req = self._active[(self._rr + off) % n]
            if req.next_row < req.n and not req.fut.cancelled():
                self._rr = (self._rr + off + 1) % max(n, 1)
                return req
        return None

    def _pinned(self, sid: int, nbytes: int):
        """The stream's pinned staging buffer, grown to at least ``nbytes``."""
        buf = self._staging[sid]
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(nbytes, 1 << 20) * 5 // 4, dtype=torch.uint8, pin_memory=True)
            self._staging[sid] = buf
        return buf

    def _upload(self, req: _Req, stream, sid: int):
        c = req.csr
        parts = [np.ascontiguousarray(c.vals, dtype=np.uint32).view(np.int32),
                 np.ascontiguousarray(c.cols, dtype=np.int32),
                 np.ascontiguousarray(c.rowptr, dtype=np.int64)]
        offs, o = [], 0
        for a in parts:
            offs.append(o)
            o += -(-a.nbytes // 256) * 256
        stage = self._pinned(sid, o)
        host = stage.numpy()
        dev = []
        with torch.cuda.stream(stream):
            for a, off in zip(parts, offs):
                host[off:off + a.nbytes] = a.view(np.uint8).reshape(-1)
                src = stage[off:off + a.nbytes].view(torch.int32 if a.dtype == np.int32
                                                     else torch.int64)
                dev.append(src.to("cuda", non_blocking=True))
            req.dev = DeviceCSR(dev[0], dev[1], dev[2], c.n_cols)
            req.idx = torch.empty((req.n, self.k), dtype=torch.int64, device="cuda")
            req.dist = torch.empty((req.n, self.k), dtype=torch.float32, device="cuda")
            req.uploaded = torch.cuda.Event()
            req.uploaded.record(stream)

    def _issue_chunk(self, req: _Req, sid: int):
        stream = self.streams[sid]
        if req.dev is None:
            self._upload(req, stream, sid)
        else:
            stream.wait_event(req.uploaded)  # chunks on other streams read the upload
        a = req.next_row
        b = min(req.n, a + self.chunk_rows)
        req.next_row = b
        req.outstanding += 1
        with torch.cuda.stream(stream):
            E = self.proj.project(req.dev, self.norm, libsize=None, col_map=self.col_map,
                                  row_start=a, n_rows=b - a, stream=stream)
            idx, dist, _ = search(E, self.E, self.refs, self.k, self.index.pool, stream=stream)
            req.idx[a:b].copy_(idx)
            req.dist[a:b].copy_(dist)
            ev = torch.cuda.Event()
            ev.record(stream)
        self._inflight.append((ev, req, "chunk", sid))

    def _issue_final(self, req: _Req, sid: int):
        stream = self.streams[sid]
        with torch.cuda.stream(stream):
            out = torch.empty((1, self.n_ref), dtype=torch.float64, device="cuda")
            buf, nbytes = workspace("mapping").get(
                _lib.lib().sh_mapping_score_workspace(self.n_ref, 1))
            _lib.call("sh_mapping_score", ptr(req.idx), ptr(req.dist), req.n, self.k, None, 1,
                      self.n_ref, 1, 0.1, self.multiplier, int(self.log_scores), ptr(out),
                      ptr(buf), nbytes, stream_handle(stream))
            req.h_idx = req.idx.to("cpu", non_blocking=True)
            req.h_dist = req.dist.to("cpu", non_blocking=True)
            req.h_scores = out.to("cpu", non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        self._inflight.append((ev, req, "final", sid))

    def _retire(self):
        keep = []
        for ev, req, kind, sid in self._inflight:
            if not ev.query():
                keep.append((ev, req, kind, sid))
                continue
            self._free_streams.append(sid)
            if kind == "chunk":
                req.outstanding -= 1
                if req.fut.cancelled() and req.outstanding == 0:
                    self._release(req)
            else:
                r = req.res
                r.indices = req.h_idx.numpy().astype(np.uint64)
                r.distances = req.h_dist.numpy()
                r.scores = req.h_scores.numpy()[0]
                r.done = time.perf_counter()
                self._release(req)
                try:
                    req.fut.set_result(r)
                except cf.InvalidStateError:  # cancelled while its last step ran
                    pass
        self._inflight = keep

    def _release(self, req: _Req):
        """Frees an admitted request's memory and slot (its work has completed)."""
        if req in self._active:
            self._used -= req.need
            self._active.remove(req)
        req.dev = req.idx = req.dist = None

    def _drop_cancelled(self):
        """Admitted requests cancelled with nothing in flight are released at once."""
        for req in [r for r in self._active if r.fut.cancelled() and r.outstanding == 0
                    and not getattr(r, "final_issued", False)]:
            self._release(req)

    def _engine(self):
        torch.cuda.set_device(0)
        while True:
            with self._lock:
                if self._stop and not self._queue and not self._active and not self._inflight:
                    return
                self._admit()
                if not self._active and not self._inflight:
                    self._lock.wait(timeout=0.05)
                    continue
            progressed = False
            self._retire()
            with self._lock:
                self._drop_cancelled()
            while self._free_streams:
                # finished requests first: their score + copy-back frees a slot
                fin = next((r for r in self._active if r.next_row >= r.n and r.outstanding == 0
                            and not r.fut.cancelled()
                            and not getattr(r, "final_issued", False)), None)
                if fin is not None:
                    fin.final_issued = True
                    self._issue_final(fin, self._free_streams.popleft())
                    progressed = True
                    continue
                req = self._pick()
                if req is None:
                    break
                self._issue_chunk(req, self._free_streams.popleft())
                progressed = True
            if not progressed:
                with self._lock:
                    if not self._stop:
                        self._lock.wait(timeout=self.poll_s)
