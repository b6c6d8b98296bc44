"""Seeded negative-binomial single-cell count generator, stored as CSR.

The model (one row = one cell, one column = one gene):

* per-gene mean      mu_g  ~ LogNormal(-1, 1.5), times a density calibration scale
* per-cell size      s_c   ~ LogNormal(0, 0.5)
* per-gene dispersion phi_g ~ U[0.1, 1]
* latent groups: each of G groups draws a LogNormal(0, 1) fold change on a random
  10% of genes (1 elsewhere); cells are assigned to groups uniformly
* optional per-gene batch shift LogNormal(0, sigma_shift) (query batches for mapping)
* count ~ NB(mean m, var m + phi m^2) = negative_binomial(1/phi, 1/(1 + phi m))

Counts are returned as CSR with u32 values, i32 columns and i64 row pointers, the
layout the HIP kernels read.  ``generate_csr`` runs on the CPU with numpy (used for
small parity cases); ``generate_csr_gpu`` draws the same model on the GPU with a
Gamma-Poisson mixture (used where a million-cell matrix must be made in seconds).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np


@dataclasses.dataclass
class CSR:
    """Row-compressed u32 counts: ``vals[rowptr[i]:rowptr[i+1]]`` at ``cols[...]``."""

    vals: np.ndarray  # uint32 [nnz]
    cols: np.ndarray  # int32 [nnz], sorted within each row
    rowptr: np.ndarray  # int64 [n_rows + 1]
    n_cols: int

    @property
    def n_rows(self) -> int:
        return int(self.rowptr.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    @property
    def shape(self):
        return (self.n_rows, self.n_cols)

    def to_dense(self, dtype=np.float64) -> np.ndarray:
        out = np.zeros((self.n_rows, self.n_cols), dtype=dtype)
        rows = np.repeat(np.arange(self.n_rows), np.diff(self.rowptr))
        out[rows, self.cols] = self.vals
        return out

    def row_slice(self, start: int, end: int) -> "CSR":
        a, b = int(self.rowptr[start]), int(self.rowptr[end])
        return CSR(
            self.vals[a:b].copy(),
            self.cols[a:b].copy(),
            (self.rowptr[start : end + 1] - a).astype(np.int64),
            self.n_cols,
        )

    def take_rows(self, rows: np.ndarray) -> "CSR":
        rows = np.asarray(rows, dtype=np.int64)
        lens = np.diff(self.rowptr)[rows]
        rp = np.zeros(rows.shape[0] + 1, dtype=np.int64)
        np.cumsum(lens, out=rp[1:])
        idx = np.concatenate(
            [np.arange(self.rowptr[r], self.rowptr[r + 1]) for r in rows]
        ) if rows.shape[0] else np.zeros(0, np.int64)
        return CSR(self.vals[idx], self.cols[idx], rp, self.n_cols)

    @staticmethod
    def from_dense(a: np.ndarray) -> "CSR":
        a = np.asarray(a)
        nzr, nzc = np.nonzero(a)
        rp = np.zeros(a.shape[0] + 1, dtype=np.int64)
        np.cumsum(np.bincount(nzr, minlength=a.shape[0]), out=rp[1:])
        return CSR(
            a[nzr, nzc].astype(np.uint32), nzc.astype(np.int32), rp, int(a.shape[1])
        )


@dataclasses.dataclass
class GeneModel:
    """Per-gene parameters shared by a reference and the query batches mapped onto it."""

    mu: np.ndarray  # float64 [d], already density-calibrated
    phi: np.ndarray  # float64 [d]
    fold: np.ndarray  # float64 [G, d]

    @property
    def n_groups(self) -> int:
        return int(self.fold.shape[0])


def _expected_density(mu, phi, s, fold_rows) -> float:
    m = s[:, None] * mu[None, :] * fold_rows
    p0 = np.power(1.0 + phi[None, :] * m, -1.0 / phi[None, :])
    return float(1.0 - p0.mean())


def make_gene_model(
    n_genes: int,
    n_groups: int = 1,
    density: float = 0.10,
    seed: int = 1000,
    fold_frac: float = 0.10,
) -> GeneModel:
    """Draws the per-gene model and scales mu so that E[nnz fraction] = ``density``."""
    rng = np.random.default_rng(seed)
    mu = rng.lognormal(-1.0, 1.5, n_genes)
    phi = rng.uniform(0.1, 1.0, n_genes)
    fold = np.ones((max(n_groups, 1), n_genes))
    if n_groups > 1:
        n_sel = max(1, int(round(fold_frac * n_genes)))
        for g in range(n_groups):
            sel = rng.choice(n_genes, n_sel, replace=False)
            fold[g, sel] = rng.lognormal(0.0, 1.0, n_sel)
    # calibrate the scale on a fixed sample of size factors and groups
    crng = np.random.default_rng(seed + 7919)
    s = crng.lognormal(0.0, 0.5, 512)
    grp = crng.integers(0, fold.shape[0], 512)
    fr = fold[grp]
    lo, hi = 1e-6, 1e6
    for _ in range(100):
        mid = math.sqrt(lo * hi)
        if _expected_density(mu * mid, phi, s, fr) < density:
            lo = mid
        else:
            hi = mid
    return GeneModel(mu * math.sqrt(lo * hi), phi, fold)


def generate_csr(
    n_cells: int,
    model: GeneModel,
    seed: int = 1000,
    batch_shift_sigma: float = 0.0,
    block_rows: int = 4096,
    return_groups: bool = False,
):
    """CPU NB draw of ``n_cells`` cells under ``model`` as CSR (numpy; small sizes)."""
    rng = np.random.default_rng(seed + 104729)
    d = model.mu.shape[0]
    shift = (
        rng.lognormal(0.0, batch_shift_sigma, d) if batch_shift_sigma > 0 else None
    )
    vals, cols, lens = [], [], []
    groups = np.empty(n_cells, dtype=np.int32)
    for start in range(0, n_cells, block_rows):
        nb = min(block_rows, n_cells - start)
        s = rng.lognormal(0.0, 0.5, nb)
        grp = rng.integers(0, model.n_groups, nb)
        groups[start : start + nb] = grp
        m = s[:, None] * model.mu[None, :] * model.fold[grp]
        if shift is not None:
            m = m * shift[None, :]
        r = 1.0 / model.phi
        p = 1.0 / (1.0 + model.phi[None, :] * m)
        c = rng.negative_binomial(np.broadcast_to(r, m.shape), p)
        nzr, nzc = np.nonzero(c)
        vals.append(c[nzr, nzc].astype(np.uint32))
        cols.append(nzc.astype(np.int32))
        lens.append(np.bincount(nzr, minlength=nb))
    rp = np.zeros(n_cells + 1, dtype=np.int64)
    if lens:
        np.cumsum(np.concatenate(lens), out=rp[1:])
    csr = CSR(
        np.concatenate(vals) if vals else np.zeros(0, np.uint32),
        np.concatenate(cols) if cols else np.zeros(0, np.int32),
        rp,
        d,
    )
    return (csr, groups) if return_groups else csr


def generate_csr_gpu(
    n_cells: int,
    model: GeneModel,
    seed: int = 1000,
    batch_shift_sigma: float = 0.0,
    block_rows: int = 50_000,
    device: str = "cuda",
):
    """GPU draw of the same model (Gamma-Poisson mixture), CSR left resident on device.

    Returns ``(vals, cols_i32, rowptr_i64)`` torch tensors on ``device``; ``vals`` is an
    int32 tensor holding the u32 count bit patterns.  Used for
    the million-cell workloads, where a numpy draw would take minutes.
    """
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(int(seed) + 15485863)
    d = model.mu.shape[0]
    mu = torch.as_tensor(model.mu, dtype=torch.float32, device=device)
    phi = torch.as_tensor(model.phi, dtype=torch.float32, device=device)
    fold = torch.as_tensor(model.fold, dtype=torch.float32, device=device)
    if batch_shift_sigma > 0:
        shift = torch.exp(
            torch.randn(d, generator=g, device=device) * batch_shift_sigma
        )
        mu = mu * shift
    vals, cols, lens = [], [], []
    for start in range(0, n_cells, block_rows):
        nb = min(block_rows, n_cells - start)
        s = torch.exp(torch.randn(nb, generator=g, device=device) * 0.5)
        grp = torch.randint(0, fold.shape[0], (nb,), generator=g, device=device)
        m = s[:, None] * mu[None, :] * fold[grp]
        shape = (1.0 / phi).expand(nb, d)
        # Gamma(shape=1/phi, scale=phi*m) then Poisson: NB(1/phi, 1/(1+phi m))
        lam = torch._standard_gamma(shape.contiguous(), generator=g) * (phi[None, :] * m)
        del m
        c = torch.poisson(lam, generator=g)
        del lam
        nz = c.nonzero()
        # counts are far below 2**31, so int32 storage has the u32 bit pattern
        vals.append(c[nz[:, 0], nz[:, 1]].to(torch.int32))
        cols.append(nz[:, 1].to(torch.int32))
        lens.append(torch.bincount(nz[:, 0], minlength=nb))
        del c, nz
    rp = torch.zeros(n_cells + 1, dtype=torch.int64, device=device)
    torch.cumsum(torch.cat(lens), 0, out=rp[1:])
    return torch.cat(vals), torch.cat(cols), rp


# Workload presets: (cells, genes, groups, density, chunk, dims, k, pool, screen dtype)
PRESETS = {
    "small": dict(n_cells=5_000, n_genes=2_000, n_groups=1, density=0.10,
                  chunk=5_000, dims=50, k=11, pool=64, screen="fp32", seed=1000),
    "100k": dict(n_cells=100_000, n_genes=2_000, n_groups=8, density=0.10,
                 chunk=20_000, dims=50, k=11, pool=64, screen="bf16", seed=1001),
    "1m": dict(n_cells=1_000_000, n_genes=2_000, n_groups=20, density=0.08,
               chunk=50_000, dims=50, k=21, pool=64, screen="bf16", seed=1002),
    "serving": dict(n_cells=500_000, n_genes=2_000, n_groups=20, density=0.08,
                    chunk=1_000, dims=50, k=11, pool=64, screen="bf16", seed=1003),
    "4m": dict(n_cells=4_000_000, n_genes=3_000, n_groups=40, density=0.06,
               chunk=100_000, dims=100, k=11, pool=128, screen="fp8", seed=1004),
}
