"""CPU oracle for scarf's graph-building path — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy and plain C, what the reference computes on the path
normalise -> mu/sigma -> incremental PCA -> projection -> kNN -> smooth_knn_dist ->
mapping scores, and the MiniBatchKMeans seed partitions beside it.  Every function cites the reference file:line it follows (paths are
relative to the scarf source tree).

It exists to check the HIP implementation in ``scarf_amd`` and is never part of it:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU baseline.  The
product path never calls into this package and fails loudly when its HIP library is
missing.

Pinning (what each part is checked against, in ``tests/test_oracle.py``):

* ``norm``    — hand-computed values on the reference's toy matrix, closed-form
                invariants (z mean 0 / var 1 with ddof=0, all-zero row -> x=0,
                clean_array fill rules).
* ``ipca``    — scikit-learn ``IncrementalPCA`` (the reference's PCA engine, pinned
                there as scikit-learn 1.6.1, run here as 1.7.2 whose ``partial_fit`` is
                unchanged) driven in scarf's block order, plus committed golden vectors
                and the single-chunk == exact-SVD special case.
* ``knn``     — numpy brute force on tiny inputs, hand-built tie and duplicate cases,
                hand-built ``fix_knn_query`` cases.
* ``smooth``  — the reference's own fixture ``knn_distances.npy`` -> ``knn_weights.npy``
                at the reference test's tolerance (1e-5), plus the zero-weight
                global-minimum rule on a hand-built case.
* ``mapping`` — hand-computed mapping scores and hand-built feature alignment.
* ``kmeans``  — scikit-learn's ``MiniBatchKMeans`` and ``kmeans_plusplus`` (the
                reference's seed-partition engine) on seeded inputs, bit for bit, plus a
                hand-computed center-update chain.
"""
