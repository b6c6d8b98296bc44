"""CPU baseline timing of the reference's algorithm (oracle only; bench.py's
cpu_baseline leg).

Reproduces scarf's CPU path on a bounded sample of a workload and extrapolates each
stage to the full cell count:
  * normalisation to a dense FP64 block per chunk (assay.py:54-64, writers.py:915-936)
    and the two-pass FP64 mean / std (graph_datastore.py:767-796)
  * scikit-learn IncrementalPCA(dims + 1) in scarf's block order (ann.py:207-256), with
    at least three real partial fits: the sample is three chunks of the workload's chunk
    size, or of a stated fraction of it (the SVD's cost per row does not depend on the
    chunk's row count once it exceeds the feature count)
  * the projection z @ loadings (ann.py:138)
  * kNN: hnswlib 0.8.0 is not installed on this image, so an exact CPU kNN
    (BLAS GEMM + argpartition); its cost per (query, reference) pair is measured on the
    sample and extrapolated quadratically to n^2 pairs.  It is reported apart from the
    linear stages, since scarf's hnswlib search grows far more slowly with n.
  * smooth_knn_dist + membership strengths, vectorised over rows (same float32
    arithmetic as ``oracle.smooth``; pinned against it in tests/test_oracle.py)

Two thread settings are timed (threadpoolctl caps BLAS): scarf's defaults --
DataStore(nthreads=2) and a single-threaded kNN unless ann_parallel
(datastore/datastore.py:55, ann.py:99-102) -- and every core of the host.
"""
from __future__ import annotations

import os
import time

import numpy as np


def smooth_knn_dist_vec(distances, k, n_iter=64, local_connectivity=1.0, bandwidth=1.0):
    """Row-vectorised form of ``oracle.smooth.smooth_knn_dist`` (float32 state)."""
    d = np.ascontiguousarray(distances, dtype=np.float32)
    n = d.shape[0]
    f32 = np.float32
    target = np.log2(k) * bandwidth
    pos = d > 0
    nnz = pos.sum(1)
    index = int(np.floor(local_connectivity))
    interp = local_connectivity - index
    # the index-th non-zero distance per row (rows sorted ascending)
    order = np.where(pos, np.arange(d.shape[1])[None, :], d.shape[1] + 1)
    srt = np.sort(order, axis=1)
    rho = np.zeros(n, dtype=f32)
    ok = nnz >= local_connectivity
    if index > 0:
        col = np.clip(srt[:, index - 1], 0, d.shape[1] - 1)
        rho[ok] = d[ok, col[ok]]
        if interp > 1e-5:
            col2 = np.clip(srt[:, index], 0, d.shape[1] - 1)
            rho[ok] += f32(interp) * (d[ok, col2[ok]] - d[ok, col[ok]])
    else:
        col = np.clip(srt[:, 0], 0, d.shape[1] - 1)
        rho[ok] = f32(interp) * d[ok, col[ok]]
    some = (~ok) & (nnz > 0)
    rho[some] = np.where(pos[some], d[some], 0).max(1)
    lo = np.zeros(n, f32)
    hi = np.full(n, np.inf, f32)
    mid = np.ones(n, f32)
    active = np.ones(n, bool)
    dd = d[:, 1:] - rho[:, None]
    for _ in range(n_iter):
        with np.errstate(over="ignore"):
            e = np.where(dd > 0, np.exp(-(dd / mid[:, None])).astype(f32), f32(1.0))
        psum = e.sum(1, dtype=f32)
        done = np.abs(psum.astype(np.float64) - target) < 1e-5
        active &= ~done
        if not active.any():
            break
        up = active & (psum > target)
        dn = active & ~(psum > target)
        hi[up] = mid[up]
        mid[up] = (lo[up] + hi[up]) / f32(2)
        lo[dn] = mid[dn]
        inf = dn & np.isinf(hi)
        mid[inf] = mid[inf] * f32(2)
        fin = dn & ~np.isinf(hi)
        mid[fin] = (lo[fin] + hi[fin]) / f32(2)
    res = mid.copy()
    rowmean = d.mean(1, dtype=np.float64).astype(f32)
    allmean = f32(d.mean(dtype=np.float64))
    fl = np.where(rho > 0, f32(1e-3) * rowmean, f32(1e-3) * allmean)
    res = np.where(res < fl, fl, res).astype(f32)
    return res, rho


def cpu_pipeline_sample(counts_dense_rows, dims: int, k: int, chunk: int, sf: float,
                        clip, n_total: int, knn_queries: int = 2000, seed: int = 7,
                        blas_threads=None, knn_threads=None):
    """Times the reference algorithm on a sample (dense uint32 rows, cut into chunks of
    ``chunk`` rows) with BLAS capped at ``blas_threads`` (kNN at ``knn_threads``) -> dict."""
    from sklearn.decomposition import IncrementalPCA
    from threadpoolctl import threadpool_limits
    from . import smooth as osmooth

    n_s, d = counts_dense_rows.shape
    t = {}
    with threadpool_limits(limits=blas_threads):
        t0 = time.perf_counter()
        blocks = []
        for s in range(0, n_s, chunk):
            c = counts_dense_rows[s:s + chunk].astype(np.float64)
            lib = c.sum(1)
            lib[lib == 0] = 1
            blocks.append(np.log1p(sf * c / lib[:, None]))
        x = np.vstack(blocks)
        mu = x.mean(0)
        sd = x.std(0)
        sd[sd == 0] = 1
        del x
        t["normalise_stats"] = time.perf_counter() - t0

        t0 = time.perf_counter()
        pca = IncrementalPCA(n_components=dims + 1)
        zb = []
        for b in blocks:
            z = (b - mu) / sd
            if clip:
                z = np.clip(z, -clip, clip)
            zb.append(z)
        order = zb[1:] + zb[:1]
        n_fits = 0
        for z in order:
            if len(z) >= dims + 1:
                pca.partial_fit(z.copy())
                n_fits += 1
        L = pca.components_[:-1].T
        t["pca_fit"] = time.perf_counter() - t0

        t0 = time.perf_counter()
        E = np.vstack([z @ L for z in zb]).astype(np.float32)
        t["project"] = time.perf_counter() - t0
        del zb, blocks

    rng = np.random.default_rng(seed)
    nq = min(knn_queries, n_s)
    qi = rng.choice(n_s, nq, replace=False)
    with threadpool_limits(limits=knn_threads):
        t0 = time.perf_counter()
        En = (E.astype(np.float64) ** 2).sum(1)
        Q = E[qi].astype(np.float64)
        dist = En[qi][:, None] + En[None, :] - 2.0 * (Q @ E.T.astype(np.float64))
        part = np.argpartition(dist, k + 1, axis=1)[:, :k + 1]
        pd = np.take_along_axis(dist, part, 1)
        o = np.argsort(pd, axis=1)
        knn_d = np.take_along_axis(pd, o, 1)[:, 1:]
        knn_i = np.take_along_axis(part, o, 1)[:, 1:]
        t_knn_sample = time.perf_counter() - t0

    with threadpool_limits(limits=blas_threads):
        t0 = time.perf_counter()
        sig, rho = smooth_knn_dist_vec(np.maximum(knn_d, 0), k, local_connectivity=1.0,
                                       bandwidth=1.5)
        _, _, w = osmooth.compute_membership_strengths(knn_i, np.maximum(knn_d, 0), sig, rho)
        t_smooth_sample = time.perf_counter() - t0

    scale = n_total / n_s
    linear = {
        "normalise_stats": t["normalise_stats"] * scale,
        "pca_fit": t["pca_fit"] * scale,
        "project": t["project"] * scale,
        "smooth": t_smooth_sample * (n_total / nq),
    }
    knn_s = t_knn_sample / (nq * n_s) * float(n_total) * float(n_total)
    lin_total = sum(linear.values())
    return {
        "cells_per_s": n_total / (lin_total + knn_s),
        "linear_stages_cells_per_s": n_total / lin_total,
        "est_seconds": dict(linear, knn_bruteforce=knn_s),
        "sample_seconds": dict(t, knn=t_knn_sample, smooth=t_smooth_sample),
        "n_sample": n_s,
        "chunk": chunk,
        "partial_fits": n_fits,
        "knn_queries": nq,
        "blas_threads": blas_threads,
        "knn_threads": knn_threads,
    }


def host_cores():
    try:
        return len(os.sched_getaffinity(0)), os.cpu_count()
    except AttributeError:  # pragma: no cover
        return os.cpu_count(), os.cpu_count()


def cpu_baseline(counts_dense_rows, dims, k, chunk, sf, clip, n_total, knn_queries=2000):
    """Both thread settings on the same sample: scarf's defaults (2 BLAS threads, 1 kNN
    thread) and every core of the host."""
    aff, ncpu = host_cores()
    ref = cpu_pipeline_sample(counts_dense_rows, dims, k, chunk, sf, clip, n_total,
                              knn_queries, blas_threads=2, knn_threads=1)
    allc = cpu_pipeline_sample(counts_dense_rows, dims, k, chunk, sf, clip, n_total,
                               knn_queries, blas_threads=aff, knn_threads=aff)
    return {"reference_defaults": ref, "all_cores": allc, "cores": aff, "cpu_count": ncpu}
