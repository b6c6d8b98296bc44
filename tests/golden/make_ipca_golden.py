"""Writes tests/golden/ipca_*.npz: scikit-learn IncrementalPCA driven in scarf's order.

scarf's PCA is scikit-learn's IncrementalPCA (third-party, pinned 1.6.1 by the
reference; run here as the installed version, whose partial_fit is unchanged), fed
block by block as AnnStream._fit_pca does (ann.py:207-256): z = (x - mu) / sigma with
the two-pass mean/std of the normalised data, the first block held back and fitted
last, and dims + 1 components with the last one dropped.  The normalisation follows
assay.py:54-64 / 812-828.  Inputs are seeded negative-binomial counts.

Run from the repository root:  python tests/golden/make_ipca_golden.py [case ...]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from sklearn.decomposition import IncrementalPCA  # noqa: E402

from scarf_amd.synth import generate_csr, make_gene_model  # noqa: E402

CASES = {
    # name: (cells, genes, groups, density, chunk, dims, sf, seed)
    "ipca_small": (600, 40, 3, 0.2, 150, 10, 1000.0, 11),
    "ipca_blocks": (1000, 120, 4, 0.1, 230, 20, 10000.0, 12),
    # wide enough (600 genes) for the GPU eigensolver's default size range
    "ipca_wide": (2000, 600, 6, 0.08, 500, 30, 10000.0, 13),
}


def scarf_order_fit(z_blocks, dims):
    pca = IncrementalPCA(n_components=dims + 1)
    reservoir, carry = None, None
    for i in z_blocks:
        if carry is not None:
            i = np.vstack((carry, i))
            carry = None
        if len(i) < dims + 1:
            carry = i
            continue
        if reservoir is None:
            reservoir = i
            continue
        pca.partial_fit(i.copy())
    last = reservoir if carry is None else np.vstack((reservoir, carry))
    pca.partial_fit(last.copy())
    return pca


def make(name):
    n, d, g, dens, chunk, dims, sf, seed = CASES[name]
    model = make_gene_model(d, g, dens, seed)
    csr = generate_csr(n, model, seed)
    counts = csr.to_dense()
    lib = counts.sum(axis=1)
    lib[lib == 0] = 1
    x = np.log1p(sf * counts / lib[:, None])
    mu = x.mean(axis=0)
    sd = x.std(axis=0)
    sd[sd == 0] = 1
    z = (x - mu) / sd
    blocks = [z[s:s + chunk] for s in range(0, n, chunk)]
    pca = scarf_order_fit(blocks, dims)
    loadings = pca.components_[:-1].T
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        vals=csr.vals, cols=csr.cols, rowptr=csr.rowptr, n_cols=np.int64(d),
        chunk=np.int64(chunk), dims=np.int64(dims), sf=np.float64(sf),
        mu=mu, sigma=sd, loadings=loadings, components=pca.components_,
        singular_values=pca.singular_values_, explained_variance_ratio=pca.explained_variance_ratio_,
        ipca_mean=pca.mean_, ipca_var=pca.var_, embedding=z @ loadings,
    )
    print("wrote", name)


if __name__ == "__main__":
    for c in (sys.argv[1:] or CASES):
        make(c)
