"""Device-side containers and plumbing: CSR counts in HBM, raw pointers, streams and
caller-owned workspaces for the C ABI.  PyTorch is used only to own device memory
and to name the current HIP stream.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


def require_gpu():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("scarf_amd runs on a ROCm GPU (torch.cuda unavailable); "
                           "there is no CPU fallback")


def ptr(t, offset_elems: int = 0) -> Optional[ctypes.c_void_p]:
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr() + offset_elems * t.element_size())


def stream_handle(stream=None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Workspace:
    """A growable byte buffer reused across calls on one stream."""

    def __init__(self, device="cuda"):
        self.device = device
        self.buf = None

    def get(self, nbytes: int):
        nbytes = max(int(nbytes), 256)
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(int(nbytes * 1.25) + 256, dtype=torch.uint8,
                                   device=self.device)
        return self.buf, self.buf.numel()


_workspaces = {}


def workspace(tag: str = "default") -> Workspace:
    key = (tag, torch.cuda.current_stream().cuda_stream)
    ws = _workspaces.get(key)
    if ws is None:
        ws = _workspaces[key] = Workspace()
    return ws


@dataclasses.dataclass
class DeviceCSR:
    """u32 counts (held as int32 bit patterns), i32 columns, i64 absolute row pointers."""

    vals: "torch.Tensor"
    cols: "torch.Tensor"
    rowptr: "torch.Tensor"
    n_cols: int

    @property
    def n_rows(self) -> int:
        return int(self.rowptr.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.vals.shape[0])

    @property
    def shape(self):
        return (self.n_rows, self.n_cols)

    def nbytes(self) -> int:
        return self.vals.numel() * 4 + self.cols.numel() * 4 + self.rowptr.numel() * 8

    @classmethod
    def from_host(cls, csr, device="cuda", validate=True) -> "DeviceCSR":
        require_gpu()
        vals = np.ascontiguousarray(csr.vals, dtype=np.uint32).view(np.int32)
        cols = np.ascontiguousarray(csr.cols, dtype=np.int32)
        rp = np.ascontiguousarray(csr.rowptr, dtype=np.int64)
        if validate:
            validate_csr_host(vals, cols, rp, csr.n_cols)
        return cls(torch.from_numpy(vals).to(device), torch.from_numpy(cols).to(device),
                   torch.from_numpy(rp).to(device), int(csr.n_cols))

    @classmethod
    def from_tensors(cls, vals, cols, rowptr, n_cols, validate=True) -> "DeviceCSR":
        out = cls(vals.contiguous(), cols.contiguous(), rowptr.contiguous(), int(n_cols))
        if validate:
            out.validate()
        return out

    def validate(self):
        """Malformed CSR (non-monotone row pointers, out-of-range columns) -> ValueError."""
        rp = self.rowptr
        if rp.dtype != torch.int64 or self.cols.dtype != torch.int32:
            raise ValueError("CSR needs int64 row pointers and int32 columns")
        if self.vals.dtype not in (torch.int32,):
            raise ValueError("CSR values must be u32 counts held in an int32 tensor")
        if rp.numel() < 1 or int(rp[0]) != 0 or int(rp[-1]) != self.vals.numel() \
                or self.cols.numel() != self.vals.numel():
            raise ValueError("CSR row pointers do not span the value array")
        if rp.numel() > 1 and bool((rp[1:] < rp[:-1]).any()):
            raise ValueError("CSR row pointers are not monotone")
        if self.cols.numel() and (bool((self.cols < 0).any()) or
                                  bool((self.cols >= self.n_cols).any())):
            raise ValueError("CSR column index out of range")
        nnz = self.cols.numel()
        if nnz > 1:
            # fast path: columns strictly increasing inside every row
            step = self.cols[1:] > self.cols[:-1]
            starts = rp[1:-1]
            starts = starts[(starts > 0) & (starts < nnz)]
            step[starts - 1] = True
            if not bool(step.all()):
                rows = torch.repeat_interleave(
                    torch.arange(rp.numel() - 1, device=rp.device), rp[1:] - rp[:-1])
                keys = rows * self.n_cols + self.cols.to(torch.int64)
                if torch.unique(keys).numel() != nnz:
                    raise ValueError("CSR has a column repeated within a row (sum duplicate "
                                     "entries before upload)")

    def to_host(self):
        from .synth import CSR
        return CSR(self.vals.cpu().numpy().view(np.uint32), self.cols.cpu().numpy(),
                   self.rowptr.cpu().numpy(), self.n_cols)


def validate_csr_host(vals, cols, rowptr, n_cols):
    if rowptr.ndim != 1 or rowptr.shape[0] < 1 or rowptr[0] != 0 or rowptr[-1] != vals.shape[0]:
        raise ValueError("CSR row pointers do not span the value array")
    if cols.shape[0] != vals.shape[0]:
        raise ValueError("CSR columns and values differ in length")
    if rowptr.shape[0] > 1 and np.any(np.diff(rowptr) < 0):
        raise ValueError("CSR row pointers are not monotone")
    if cols.shape[0] and (cols.min() < 0 or cols.max() >= n_cols):
        raise ValueError("CSR column index out of range")
    nnz = cols.shape[0]
    if nnz > 1:
        # fast path: columns strictly increasing inside every row
        step = cols[1:] > cols[:-1]
        starts = np.asarray(rowptr[1:-1])
        starts = starts[(starts > 0) & (starts < nnz)]
        step[starts - 1] = True
        if not step.all():
            rows = np.repeat(np.arange(rowptr.shape[0] - 1, dtype=np.int64), np.diff(rowptr))
            keys = rows * int(n_cols) + cols.astype(np.int64)
            if np.unique(keys).shape[0] != nnz:
                raise ValueError("CSR has a column repeated within a row (sum duplicate "
                                 "entries before upload)")


def col_map_tensor(n_cols_in: int, feat_idx=None, device="cuda"):
    """Input column -> rank among the (sorted) selected features, or -1.  None = all."""
    if feat_idx is None:
        return None, n_cols_in
    feat_idx = np.asarray(feat_idx)
    if feat_idx.dtype == bool:
        if feat_idx.shape[0] != n_cols_in:
            raise ValueError("boolean feature mask length differs from the column count")
        feat_idx = np.where(feat_idx)[0]
    feat_idx = np.asarray(feat_idx, dtype=np.int64)
    if np.unique(feat_idx).shape[0] != feat_idx.shape[0]:
        raise ValueError("feature subset has duplicate indices")
    if feat_idx.size and (feat_idx.min() < 0 or feat_idx.max() >= n_cols_in):
        raise ValueError("feature index out of range")
    cm = np.full(n_cols_in, -1, dtype=np.int32)
    order = np.sort(feat_idx)
    cm[order] = np.arange(order.shape[0], dtype=np.int32)
    return torch.from_numpy(cm).to(device), int(order.shape[0])


def col_map_from_positions(n_cols_in: int, positions: np.ndarray, device="cuda"):
    """Explicit map: positions[c] = output feature of input column c, or -1."""
    cm = np.asarray(positions, dtype=np.int32)
    if cm.shape[0] != n_cols_in:
        raise ValueError("column map length differs from the column count")
    used = cm[cm >= 0]
    if np.unique(used).shape[0] != used.shape[0]:
        raise ValueError("column map sends two columns to one feature")
    return torch.from_numpy(cm).to(device)
