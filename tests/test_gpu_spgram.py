"""GPU: the sparse FP64 Gram (sh_sparse_gram) against FP64 Grams of z formed on the host from
the oracle's normalisation, and against the dense GPU path (sh_densify + sh_gram).

The bound is the one of the dense FP64 test: |G - G_ref| <= 1e-12 * sum_r |z_ri z_rj| for
every entry.  The cases cover row subsets with repeats, all-zero rows, rows denser than 64
entries per 128-feature block, a feature subset, clipping, no log transform, feature counts
on both sides of a block edge (the ones column alone in its block) and 24 blocks."""
import sys

import numpy as np
import pytest

from oracle import norm as onorm
from scarf_amd.synth import CSR, generate_csr, make_gene_model

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _counts(n, d, seed, zero_rows=0, dense_rows=0):
    model = make_gene_model(d, 4, 0.1, seed)
    dense = generate_csr(n, model, seed).to_dense().astype(np.uint32)
    rng = np.random.default_rng(seed + 1)
    if zero_rows:
        dense[rng.choice(n, zero_rows, replace=False)] = 0
    if dense_rows:
        dense[rng.choice(n, dense_rows, replace=False)] = rng.integers(1, 40, (dense_rows, d))
    return CSR.from_dense(dense), dense


def _host_gram(dense, st, p, feat, rows, disable_scaling=False):
    x = onorm.normed(dense.astype(np.float64), sf=p.size_factor, log_transform=p.log_transform,
                     feat_idx=feat)
    mu = st.mu.cpu().numpy()
    sigma = st.sigma.cpu().numpy()
    if disable_scaling:
        z = x - 0.0
    else:
        z = (x - mu) * (1.0 / sigma)
    if p.clip:
        z = np.clip(z, -p.clip, p.clip)
    z = z[rows]
    return z.T @ z, z.sum(0), np.abs(z).T @ np.abs(z)


CASES = {
    # name: (n, d, zero_rows, dense_rows, n_feat, rows, clip, log)
    "rows_100_2900": (3000, 300, 60, 0, None, "range", None, True),
    "clip_2": (3000, 300, 60, 0, None, "range", 2.0, True),
    "subset_repeats": (3000, 300, 30, 0, 220, "repeats", None, True),
    "dense_rows": (2000, 300, 20, 40, None, "all", None, True),
    "no_log": (2000, 300, 0, 10, None, "all", None, False),
    "d255": (2500, 255, 10, 5, None, "all", None, True),
    "d256": (2500, 256, 10, 5, None, "all", 3.0, True),
    "d3000": (4000, 3000, 10, 3, None, "all", None, True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_sparse_gram_vs_host(cuda, name):
    import torch
    from scarf_amd.device import DeviceCSR, col_map_tensor
    from scarf_amd.normalize import NormParams, feature_stats
    from scarf_amd.pca import gram_of_rows
    n, d, zr, dr, nf, rsel, clip, log_t = CASES[name]
    say("sparse gram case", name, n, d)
    csr, dense = _counts(n, d, 7 + d, zr, dr)
    rng = np.random.default_rng(3)
    feat = np.sort(rng.choice(d, nf, replace=False)) if nf else None
    if rsel == "range":
        rows = np.arange(100, n - 100)
    elif rsel == "repeats":
        rows = rng.integers(0, n, n + 500)  # unsorted, with repeats
    else:
        rows = np.arange(n)
    dev = DeviceCSR.from_host(csr)
    cm, dd = col_map_tensor(csr.n_cols, feat)
    p = NormParams(size_factor=1000.0, log_transform=log_t, clip=clip)
    st = feature_stats(dev, 1000, p, col_map=cm, d=dd)
    rows_t = torch.from_numpy(rows.astype(np.int64)).cuda()
    G, cs, m = gram_of_rows(dev, st, p, rows_t, col_map=cm, method="sparse")
    Gd, csd, _ = gram_of_rows(dev, st, p, rows_t, col_map=cm, method="dense")
    Gr, csr_, scale = _host_gram(dense, st, p, feat, rows)
    assert m == rows.shape[0]
    den = np.maximum(scale, 1e-300)  # a feature with no counts: exact zeros required
    err = (np.abs(G - Gr) / den).max()
    err_d = (np.abs(G - Gd) / den).max()
    say("  max error / sum|z z|: vs host", err, "vs dense GPU", err_d)
    assert err < 1e-12 and err_d < 1e-12
    assert np.array_equal(G, G.T)
    tol = 1e-12 * np.abs(scale.diagonal()).max()
    np.testing.assert_allclose(cs, csr_, rtol=0, atol=tol)
    np.testing.assert_allclose(cs, csd, rtol=0, atol=tol)


def test_sparse_gram_bit_reproducible_and_accumulates(cuda):
    """Integer accumulation: repeated calls give identical bits; accumulate=1 adds (2G is
    exact in FP64); no rows gives zeros."""
    import torch
    from scarf_amd import _lib
    from scarf_amd.device import DeviceCSR, ptr
    from scarf_amd.normalize import NormParams, feature_stats
    from scarf_amd.pca import _z_params, gram_of_rows
    csr, dense = _counts(6000, 400, 11, 50, 20)
    dev = DeviceCSR.from_host(csr)
    p = NormParams()
    st = feature_stats(dev, 2000, p)
    rows = torch.arange(6000, device="cuda", dtype=torch.int64)
    say("sparse gram repeat / accumulate")
    G1, c1, _ = gram_of_rows(dev, st, p, rows, method="sparse")
    G2, c2, _ = gram_of_rows(dev, st, p, rows, method="sparse")
    assert np.array_equal(G1, G2) and np.array_equal(c1, c2)
    d = st.d
    G = torch.from_numpy(G1).cuda()
    cs = torch.from_numpy(c1).cuda()
    center, scale = _z_params(st, False)
    nnz = int(dev.rowptr[-1].item())
    nbytes = _lib.lib().sh_sparse_gram_workspace(6000, nnz, d)
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    _lib.call("sh_sparse_gram", ptr(dev.vals), ptr(dev.cols), ptr(dev.rowptr), ptr(rows), 6000,
              nnz, None, d, d, float(p.size_factor), 1, ptr(st.libsize), ptr(center),
              ptr(scale), 0.0, 1, ptr(G), ptr(cs), ptr(ws), nbytes, 0)
    torch.cuda.synchronize()
    assert np.array_equal(G.cpu().numpy(), 2.0 * G1)
    assert np.array_equal(cs.cpu().numpy(), 2.0 * c1)
    # a workspace sized for too few entries is refused with a message, not overrun
    small = _lib.lib().sh_sparse_gram_workspace(6000, nnz // 2, d)
    with pytest.raises(ValueError, match="stored entries"):
        _lib.call("sh_sparse_gram", ptr(dev.vals), ptr(dev.cols), ptr(dev.rowptr), ptr(rows),
                  6000, nnz // 2, None, d, d, float(p.size_factor), 1, ptr(st.libsize),
                  ptr(center), ptr(scale), 0.0, 0, ptr(G), ptr(cs), ptr(ws), small, 0)
    empty = torch.zeros(0, device="cuda", dtype=torch.int64)
    G0, c0, m0 = gram_of_rows(dev, st, p, empty, method="sparse")
    assert m0 == 0 and not G0.any() and not c0.any()


def test_gram_method_choice(cuda):
    """method="auto" takes the dense MFMA form above 30% density and the sparse one below;
    both agree with the host Gram.  The sparse form refuses more than 8063 features."""
    import torch
    from scarf_amd.device import DeviceCSR
    from scarf_amd.normalize import NormParams, feature_stats
    from scarf_amd.pca import SPARSE_GRAM_MAX_D, gram_of_rows
    rng = np.random.default_rng(5)
    dense = rng.poisson(1.5, (1500, 200)).astype(np.uint32)  # ~78% non-zero
    csr = CSR.from_dense(dense)
    dev = DeviceCSR.from_host(csr)
    p = NormParams()
    st = feature_stats(dev, 500, p)
    rows = torch.arange(1500, device="cuda", dtype=torch.int64)
    say("gram method choice, density", csr.vals.shape[0] / dense.size)
    G_auto, cs_auto, _ = gram_of_rows(dev, st, p, rows)
    G_dense, cs_dense, _ = gram_of_rows(dev, st, p, rows, method="dense")
    G_sparse, _, _ = gram_of_rows(dev, st, p, rows, method="sparse")
    assert np.array_equal(G_auto, G_dense) and np.array_equal(cs_auto, cs_dense)
    Gr, _, scale = _host_gram(dense, st, p, None, np.arange(1500))
    den = np.maximum(scale, 1e-300)
    assert (np.abs(G_sparse - Gr) / den).max() < 1e-12
    assert (np.abs(G_dense - Gr) / den).max() < 1e-12
    wide = CSR.from_dense(np.eye(4, SPARSE_GRAM_MAX_D + 1, dtype=np.uint32))
    dw = DeviceCSR.from_host(wide)
    stw = feature_stats(dw, 4, p)
    with pytest.raises(ValueError, match="features"):
        gram_of_rows(dw, stw, p, torch.arange(4, device="cuda", dtype=torch.int64),
                     method="sparse")
