"""Writes tests/golden/config_{small,100k}.npz: scikit-learn IncrementalPCA on the
small (5,000 x 2,000) and 100k (100,000 x 2,000) workloads, driven in scarf's order.

The counts are not stored: both sides regenerate them with the seeded CPU generator
(``scarf_amd.synth.generate_csr``, numpy's default_rng), and the fixture keeps a
checksum of the matrix (nnz, sum of values, sum of column indices, sum of row
pointers) so a test can tell that it is looking at the same input.

Semantics of the workloads: library size over all features, x = log1p(1e4 c / l)
(assay.py:54-64, 812-828), two-pass mean and ddof-0 std with clean_array
(graph_datastore.py:767-796, utils.py:143-154), z = clip((x - mu) / sigma, -10, 10),
IncrementalPCA(dims + 1) fed block by block with the first block held back and
fitted last (ann.py:207-256), loadings = components_[:-1].T.  Stored: mu, sigma, the
loadings, singular values, explained variance ratios and the PCA scores of a fixed
sample of rows.

Run from the repository root:  python tests/golden/make_config_golden.py [case ...]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from sklearn.decomposition import IncrementalPCA  # noqa: E402

from scarf_amd.synth import generate_csr, make_gene_model  # noqa: E402

CASES = {
    # name: (cells, genes, groups, density, chunk, dims, sf, clip, seed, score rows)
    "config_small": (5_000, 2_000, 1, 0.10, 5_000, 50, 1e4, 10.0, 1000, 1000),
    "config_100k": (100_000, 2_000, 8, 0.10, 20_000, 50, 1e4, 10.0, 1001, 2000),
}


def checksum(csr):
    return np.array([csr.nnz, int(csr.vals.astype(np.uint64).sum()),
                     int(csr.cols.astype(np.int64).sum()), int(csr.rowptr.sum())],
                    dtype=np.int64)


def make(name):
    n, d, g, dens, chunk, dims, sf, clip, seed, n_score = CASES[name]
    model = make_gene_model(d, g, dens, seed)
    csr = generate_csr(n, model, seed)
    print(name, "nnz", csr.nnz, flush=True)
    # two passes over row blocks (a dense 100k x 2000 FP64 copy of x is 1.6 GB)
    row_of = np.repeat(np.arange(n), np.diff(csr.rowptr))
    lib = np.bincount(row_of, weights=csr.vals.astype(np.float64), minlength=n)
    lib[lib == 0] = 1.0

    def x_rows(a, b):
        sub = csr.row_slice(a, b)
        c = sub.to_dense()
        return np.log1p(sf * c / lib[a:b, None])

    blocks_x = [x_rows(s, min(n, s + chunk)) for s in range(0, n, chunk)]
    x = np.vstack(blocks_x)
    mu = x.mean(axis=0)
    sd = x.std(axis=0)
    sd[sd == 0] = 1.0
    z = np.clip((x - mu) / sd, -clip, clip)
    del x, blocks_x
    blocks = [z[s:s + chunk] for s in range(0, n, chunk)]
    pca = IncrementalPCA(n_components=dims + 1)
    reservoir, carry = None, None
    for i in blocks:
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
    loadings = pca.components_[:-1].T
    rows = np.sort(np.random.default_rng(7).choice(n, n_score, replace=False))
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        checksum=checksum(csr), n_cells=np.int64(n), n_genes=np.int64(d), groups=np.int64(g),
        density=np.float64(dens), seed=np.int64(seed), chunk=np.int64(chunk),
        dims=np.int64(dims), sf=np.float64(sf), clip=np.float64(clip),
        mu=mu, sigma=sd, loadings=loadings, singular_values=pca.singular_values_,
        explained_variance_ratio=pca.explained_variance_ratio_, score_rows=rows,
        scores=z[rows] @ loadings,
    )
    print("wrote", name, flush=True)


if __name__ == "__main__":
    for c in (sys.argv[1:] or CASES):
        make(c)
