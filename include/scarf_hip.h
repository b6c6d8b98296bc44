/* scarf_hip.h — C ABI of libscarfhip.so, the MI355X (gfx950) kernels behind scarf's
 * graph-building path (normalise -> mu/sigma -> PCA -> projection -> kNN ->
 * smooth_knn_dist -> mapping scores).
 *
 * Conventions
 * -----------
 * - All array arguments are DEVICE pointers owned by the caller; sizes are element
 *   counts.  Nothing here allocates device memory: entry points that need scratch
 *   take a caller-provided workspace whose size the matching *_workspace() call gives.
 * - Every call is ordered on the HIP stream passed as `stream` (a hipStream_t; NULL
 *   means the default stream) and returns SH_OK (0) or a negative code; the message
 *   of the last failure on the calling thread is available from sh_last_error().
 * - Count matrices are CSR: u32 values, i32 column ids (sorted within a row), i64 row
 *   pointers holding ABSOLUTE offsets into the value/column arrays, so a row chunk is
 *   addressed by offsetting `rowptr` alone.
 * - `col_map` (length n_cols_in, may be NULL for the identity) maps an input column to
 *   its position among the d selected features, or -1 to ignore the column.  This is
 *   the feature subset of scarf's `feat_key` and the feature alignment of mapping.
 *
 * The reference (scarf 0.32.3) has no native interface: its path is Python calling
 * numpy/dask, scikit-learn IncrementalPCA, hnswlib and umap-learn.  Each entry point
 * below names the reference code it replaces (paths relative to the scarf tree).
 */
#ifndef SCARF_HIP_H
#define SCARF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SH_OK 0
#define SH_EINVAL (-1)      /* bad argument: shape, pointer, size, malformed CSR */
#define SH_EHIP (-2)        /* HIP runtime error */
#define SH_EWORKSPACE (-3)  /* workspace smaller than *_workspace() asked for */
#define SH_EUNSUPPORTED (-4)

/* Screening arithmetic for sh_knn_pack / sh_knn_screen. */
#define SH_F32 0
#define SH_BF16 1
#define SH_FP8 2
/* Row type of the k-means entry points (SH_F32 rows are promoted to FP64 exactly). */
#define SH_F64 3

/* Bits of the per-query flags written by sh_knn_rerank / sh_knn_exact. */
#define SH_FLAG_UNCERTIFIED 1 /* FP32 screen could not prove its top-k exact */
#define SH_FLAG_SELF_MISSING 2 /* fix_knn_query had to drop the last neighbour */
#define SH_FLAG_SELF_NOT_FIRST 4 /* column 0 was not the query (a miss, ann.py:34-35) */
#define SH_FLAG_THIN 8 /* low-precision screen: k-th distance within margin x the pool's
                          largest screening error of the pool's cut-off (re-screen it) */

typedef void *sh_stream_t;

int sh_version(void);
const char *sh_last_error(void);

/* ---- normalisation and feature statistics ------------------------------------
 * Per-cell library size over the mapped columns (0 -> 1) unless `libsize_in` is
 * given, x = log1p(sf * c / libsize) (or sf * c / libsize when log_transform == 0),
 * and per-chunk FP64 sums of x and x^2 for every feature (implicit zeros add 0):
 * rows [0, n_rows) are split into chunks of `chunk_rows` (<= 0: one chunk) and
 * sum_out / sumsq_out are [n_chunks x d].  All chunks run in one launch; the
 * accumulation order is fixed, so results are deterministic.
 * Replaces: assay.py:812-828 (RNAassay.normed with renormalize_subset),
 *           assay.py:41-64 (norm_lib_size[_log]),
 *           graph_datastore.py:767-796 (data.mean(0) / data.std(0), per chunk). */
int sh_norm_stats(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
                  int64_t n_rows, int64_t chunk_rows, const int32_t *col_map,
                  int32_t n_cols_in, int32_t d, double sf, int log_transform,
                  const double *libsize_in, double *libsize_out, double *sum_out,
                  double *sumsq_out, void *workspace, size_t workspace_bytes,
                  sh_stream_t stream);
size_t sh_norm_stats_workspace(int64_t n_rows, int64_t chunk_rows, int32_t d);

/* Merges per-chunk (count, sum, sumsq) [n_chunks][d] with Chan's pairwise formula
 * into mu and sigma (population, ddof=0), then applies clean_array: NaN/inf -> 0,
 * mu 0 stays 0, sigma 0 -> 1.  Optional per-chunk means [n_chunks][d].
 * Replaces: graph_datastore.py:767-796 + utils.py:143-154 (clean_array). */
int sh_stats_merge(const double *sums, const double *sumsqs, const int64_t *counts,
                   int32_t n_chunks, int32_t d, double *mu_out, double *sigma_out,
                   double *chunk_mean_out, sh_stream_t stream);

/* Dense tile of z = (x - center) * scale (clipped to +-clip when clip > 0), FP64
 * (out_f64 != 0) or FP32, for the rows `rows[0..n_rows)` (NULL: rows 0..n_rows-1 of
 * the CSR).  Padding columns: column d holds 1.0 when ld > d (so sh_gram's ones_col
 * yields the column sums), columns d+1..ld-1 hold 0.  Feeds sh_gram.
 * Replaces: ann.py:191-192 (transform_z) on the blocks fed to
 *           IncrementalPCA.partial_fit (ann.py:229-241). */
int sh_densify(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
               const int64_t *rows, int64_t n_rows, const int32_t *col_map,
               int32_t n_cols_in, int32_t d, double sf, int log_transform,
               const double *libsize, const double *center, const double *scale,
               double clip, int out_f64, void *out, int64_t ld, sh_stream_t stream);

/* G = Z^T Z (FP64 [d x d], full symmetric) of a dense Z [n x ld] on MFMA: FP64 Z on
 * v_mfma_f64_16x16x4f64 (z_f64 != 0), or FP32 Z on v_mfma_f32_32x32x2f32 with FP32
 * accumulation per row slab; slabs are summed in FP64 in a fixed order.  Column sums
 * of Z go to colsum_out [d] (may be NULL): with ones_col != 0 (Z[:, d] == 1, as
 * sh_densify writes) they are read off the Gram's column d at no extra pass.
 * accumulate != 0 adds to G_out / colsum_out instead of overwriting them.  ld must be
 * a multiple of 128; columns past d must be as sh_densify writes them.
 * Replaces: the SVD inside IncrementalPCA.partial_fit (_incremental_pca.py:345-363,
 *           called from ann.py:241,250): right singular vectors of the augmented
 *           chunk are eigenvectors of its Gram matrix. */
int sh_gram(const void *Z, int z_f64, int64_t n, int32_t d, int64_t ld, int ones_col,
            int accumulate, double *G_out, double *colsum_out, void *workspace,
            size_t workspace_bytes, sh_stream_t stream);
size_t sh_gram_workspace(int64_t n, int32_t d, int z_f64);

/* The same G = Z^T Z and column sums as sh_densify + sh_gram (z = clip((x - center) *
 * scale) over rows[0..n_rows), NULL: rows 0..n_rows-1), computed from the CSR entries
 * without densifying z: with c_f the z of a zero count and delta = z - c non-zero only
 * at stored entries, Z^T Z = delta^T delta + c s^T + s c^T + n c c^T (s = delta^T 1).
 * delta^T delta is summed exactly in int64 fixed point (per-feature power-of-two scales
 * from the column maxima of |delta|, products rounded once to 2^-F of those scales with
 * F <= 50 set by the row-slab size), so the result is bit-reproducible; the slab sums and
 * the rank-one terms are added in FP64 in a fixed order.  `nnz` bounds the stored
 * entries of the selected rows (it sizes the packed entries; the call checks it against
 * the rows' true count, which costs one stream synchronisation).  d + 1 <= 8064; col_map
 * must be injective, as for sh_densify.  A non-finite delta turns its feature's row and
 * column of G into NaN.  accumulate != 0 adds to G_out / colsum_out.
 * Replaces: the Gram of the augmented chunk whose SVD IncrementalPCA.partial_fit takes
 *           (_incremental_pca.py:345-363, called from ann.py:241,250), on z of
 *           ann.py:191-192. */
int sh_sparse_gram(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
                   const int64_t *rows, int64_t n_rows, int64_t nnz, const int32_t *col_map,
                   int32_t n_cols_in, int32_t d, double sf, int log_transform,
                   const double *libsize, const double *center, const double *scale,
                   double clip, int accumulate, double *G_out, double *colsum_out,
                   void *workspace, size_t workspace_bytes, sh_stream_t stream);
size_t sh_sparse_gram_workspace(int64_t n_rows, int64_t nnz, int32_t d);

/* ---- symmetric eigensolve of the PCA merge matrix (FP64) --------------------------
 * The top (dims + 1) eigenpairs of the d x d merge matrix replace the SVD of the
 * augmented chunk in IncrementalPCA.partial_fit (_incremental_pca.py:345-363, called
 * from ann.py:241,250).  The d-sized work runs here; the caller solves the small
 * block-tridiagonal problem and tests convergence.
 *
 * sh_merge_matrix: M = (G - n bm bm^T) + Vt^T diag(s2) Vt + mc mc^T, all FP64 d x d
 * row-major (G symmetric); Vt [nc x d] may be NULL (first fit: no previous
 * components), mc may be NULL.  Restates _incremental_pca.py:335-356 (the chunk
 * centring, the previous components scaled by their singular values and the
 * mean-correction row), formed as a Gram instead of stacked rows. */
int sh_merge_matrix(const double *G, int32_t d, double n, const double *bm,
                    const double *Vt, const double *s2, int32_t nc, const double *mc,
                    double *M, sh_stream_t stream);

/* Block Lanczos (block size 8) with full two-pass classical Gram-Schmidt and
 * CholeskyQR2: runs block steps j0 .. j0 + n_steps - 1 on M (d x d symmetric).
 * V [kmax x d] holds the orthonormal basis, one vector per row; rows below 8 (j0 + 1)
 * come from the caller (the starting block, or after a thick restart the kept Ritz
 * vectors followed by the newest block).  Step j orthogonalises M V_j against every
 * row below 8 (j + 1) and writes rows 8(j+1)..8(j+1)+7, the diagonal block
 * T_diag[j] = V_j M V_j^T (8 x 8) and T_sub[j] = R_j (8 x 8 upper triangular, the block
 * below the diagonal of the Lanczos matrix).  status (i32, set to 0 by the caller)
 * becomes 1 when a new block is numerically rank deficient (basis exhausted).
 * Needs 8 (j0 + n_steps + 1) <= kmax and kmax % 8 == 0. */
size_t sh_lanczos_workspace(int32_t d, int32_t kmax);
int sh_lanczos_steps(const double *M, int32_t d, double *V, int32_t kmax, int32_t j0,
                     int32_t n_steps, double *T_diag, double *T_sub, int32_t *status,
                     void *workspace, size_t workspace_bytes, sh_stream_t stream);

/* FP64 GEMM on MFMA: C[m x n] = alpha op(A) B + beta C, op(A) = A (m x k, transA 0)
 * or A^T (A stored k x m, transA 1), row-major.  Used for the Ritz vectors
 * (S^T V) and their residual check (Vt M). */
int sh_dgemm(int transA, int64_t m, int64_t n, int64_t k, double alpha, const double *A,
             int64_t lda, const double *B, int64_t ldb, double beta, double *C, int64_t ldc,
             sh_stream_t stream);

/* Block subspace iteration with Chebyshev filtering, the default GPU eigensolver of the
 * same merge matrix (same replaced call as above).  A block of vectors is stored as the
 * first ncols columns of a d x ldy row-major array (ncols and ldy multiples of 16; ncols
 * split into equal groups of at most 128 columns, i.e. ncols <= 128 or a multiple of 32
 * up to 256, ...).
 * sh_symm_block: one filter step Y2 = alpha M Y1 + beta Y1 + gamma Y0 with M d x d
 * symmetric (Y0 may be NULL when gamma == 0; Y2 must not alias Y1 or Y0; Y1 16-byte
 * aligned).  The product is summed over k slices in a fixed order: bit-reproducible.
 * sh_block_gram: C [na x nb] (row-major, ld nb) = A^T B for column blocks A (d x lda)
 * and B (d x ldb): the Gram of a block (CholeskyQR) and the projected matrix X^T M X.
 * sh_block_residual: res[c] = ||XM[:, c] - theta[c] X[:, c]|| for c < b (Ritz residuals). */
size_t sh_block_gram_workspace(int32_t d, int32_t na, int32_t nb);
int sh_block_gram(const double *A, int64_t lda, int32_t na, const double *B, int64_t ldb,
                  int32_t nb, int32_t d, double *C, void *workspace, size_t workspace_bytes,
                  sh_stream_t stream);
size_t sh_symm_block_workspace(int32_t d, int32_t ncols, int64_t ldy);
int sh_symm_block(const double *M, int32_t d, const double *Y1, const double *Y0, int32_t ncols,
                  int64_t ldy, double alpha, double beta, double gamma, double *Y2,
                  void *workspace, size_t workspace_bytes, sh_stream_t stream);
int sh_block_residual(const double *X, const double *XM, int32_t d, int32_t b, int64_t ldy,
                      const double *theta, double *res, sh_stream_t stream);

/* Dense blocks of normalised values (x FP32 or FP64 [n x ldx], SH_F32 / SH_F64), as
 * scarf's callers hand them to AnnStream.  z = (x - center) / sigma with an IEEE division
 * (center and sigma FP64 [d], both NULL: no scaling), clipped to +-clip when clip > 0.
 * sh_dense_z writes z (FP64 [n x ldz]); sh_dense_project writes z . W (W FP64 [d x D],
 * out FP64 [n x ldo], D <= 128), features summed in increasing order per row.
 * Replaces: ann.py:191-192 (transform_z) and ann.py:126-138 (reducer on a numpy block),
 *           as called at mapping_datastore.py:204, knn_utils.py:42, metrics.py:312-315. */
int sh_dense_z(const void *x, int x_dtype, int64_t n, int32_t d, int64_t ldx,
               const double *center, const double *sigma, double clip, double *z, int64_t ldz,
               sh_stream_t stream);
int sh_dense_project(const void *x, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                     const double *center, const double *sigma, double clip, const double *W,
                     int32_t D, double *out, int64_t ldo, sh_stream_t stream);

/* Sparse z-score projection: E[r] = sum_f z_rf W[f] with z = (x - mu) * inv_sigma
 * (clipped to +-clip when clip > 0), evaluated as c + sum over the row's non-zeros of
 * (z(x) - z(0)) * W[f], where c = sum_f z(0)_f W[f] (+ alignment fill, see mapping)
 * is precomputed by the caller.  W is FP32 [d x ldw] with ldw a multiple of 4 and W
 * 16-byte aligned (rows are gathered 16 bytes at a time), E is FP32 [n_rows x lde].
 * libsize NULL: computed over the mapped columns (mapping's target normalisation).
 * Replaces: ann.py:129-138 (reducer = transform_z(x).dot(loadings)),
 *           mapping_datastore.py:198-204 (target projection). */
int sh_project(const uint32_t *vals, const int32_t *cols, const int64_t *rowptr,
               int64_t n_rows, const int32_t *col_map, int32_t n_cols_in, int32_t d,
               double sf, int log_transform, const double *libsize, const float *mu,
               const float *inv_sigma, float clip, const float *W, int64_t ldw,
               int32_t D, const float *c, float *E_out, int64_t lde, sh_stream_t stream);

/* ---- k nearest neighbours (squared L2) ------------------------------------------
 * Replaces hnswlib.Index (ann.py:14-28, 194-205, 290-326) with an exact screen:
 *   1. sh_knn_pack       embedding rows -> MFMA operand layout (F32 / BF16 / FP8)
 *   2. sh_knn_screen     ||r||^2 - 2 q.r tiles on MFMA; per (query, reference slice)
 *                        the `pool` smallest (screened distance, index) pairs
 *   3. sh_knn_rerank     merge of the slices, exact FP64 distances of the pool,
 *                        top-k by (distance, lower index), fix_knn_query self repair
 *                        (ann.py:31-52), and an error-bound certificate
 *   4. sh_knn_exact      FP64 brute force for rows the certificate rejects (and the
 *                        FP64 reference for recall measurements). */
int32_t sh_knn_pack_k(int dtype, int32_t D);
/* Bytes per packed row: K elements (F32 4 B, BF16 2 B), or for FP8 K bytes of E4M3
 * followed by a 16-byte trailer {f32 ||r||^2, i32 E8M0 scale exponent + 127, pad}. */
int64_t sh_knn_pack_row_bytes(int dtype, int32_t D);
int sh_knn_pack(const float *E, int64_t n, int32_t D, int64_t lde, int dtype, int role,
                void *out, float *aux_out, sh_stream_t stream);
/* seed_refs (NULL, or one reference row per query): each block of 256 consecutive
 * queries first screens the block's seed rows and starts its threshold at their pool-th
 * smallest key, which skips candidates but keeps the same pools.  The seed rows of a
 * block must be distinct; rows outside a slice are ignored for it.  Self queries in
 * locality order pass their own reference rows. */
int sh_knn_screen(const void *qpack, int64_t nq, const void *rpack, int64_t nr,
                  int32_t D, int dtype, const float *q_aux, const float *r_aux,
                  int32_t pool, int32_t n_slices, const int64_t *seed_refs,
                  uint64_t *cand_keys, int32_t *cand_cnt, void *workspace,
                  size_t workspace_bytes, sh_stream_t stream);
size_t sh_knn_screen_workspace(int64_t nq, int32_t pool, int32_t n_slices);
/* Self queries: query q is reference self_idx[q] (self_idx non-NULL) or
 * self_offset + q (self_offset >= 0); k + 1 neighbours are ranked and the query's
 * own column removed as fix_knn_query does.  Neither: a plain k-neighbour query.
 * err_coef >= 0: the FP32 certificate (SH_FLAG_UNCERTIFIED).  margin > 0: the margin
 * test of a low-precision screen (SH_FLAG_THIN): the k-th exact distance must lie more
 * than margin x max over the pool |screened key + ||q||^2 - exact distance| below the
 * pool's cut-off key + ||q||^2. */
int sh_knn_rerank(const float *Q, int64_t nq, int64_t ldq, const float *R, int64_t nr,
                  int64_t ldr, int32_t D, const uint64_t *cand_keys,
                  const int32_t *cand_cnt, int32_t pool, int32_t n_slices,
                  int32_t k, int64_t self_offset, const int64_t *self_idx,
                  double err_coef, float r_norm_max, double margin, int64_t *idx_out,
                  float *dist_out, int32_t *flags_out, sh_stream_t stream);
/* FP64 brute force for the listed query rows (all rows when q_rows is NULL).  flags_out
 * is read and rewritten: an SH_FLAG_UNCERTIFIED bit already set for a row (by the
 * re-rank whose certificate failed) is kept, the self bits are replaced. */
int sh_knn_exact(const float *Q, int64_t nq, int64_t ldq, const int64_t *q_rows,
                 int64_t n_sel, const float *R, int64_t nr, int64_t ldr, int32_t D,
                 int32_t k, int64_t self_offset, const int64_t *self_idx,
                 int64_t *idx_out, float *dist_out, double *dist64_out,
                 int32_t *flags_out, sh_stream_t stream);
/* fix_knn_query on an existing (k+1)-column table [n x k1]: removes each row's own
 * index (ref_idx[row]) or, if absent, its last column.  Outputs [n x (k1-1)].
 * Replaces: ann.py:31-52 (fix_knn_query). */
int sh_fix_knn_query(const int64_t *idx_in, const float *dist_in, const int64_t *ref_idx,
                     int64_t n, int32_t k1, int64_t *idx_out, float *dist_out,
                     int32_t *flags_out, sh_stream_t stream);

/* Locality order of embedding rows: a stable sort of the Morton code of the first three
 * coordinates (10 bits each over their range).  perm_out [n] (i64) lists the rows in that
 * order.  The screen takes its queries in this order so that the 32 queries of a tile are
 * neighbours; candidate pools do not depend on query order, so results are unchanged.
 * No counterpart in the reference (hnswlib queries one row at a time). */
size_t sh_locality_order_workspace(int64_t n);
int sh_locality_order(const float *E, int64_t n, int32_t D, int64_t lde, int64_t *perm_out,
                      void *workspace, size_t workspace_bytes, sh_stream_t stream);
/* Row permutation in 4-byte words: scatter == 0: dst[i] = src[perm[i]];
 * scatter != 0: dst[perm[i]] = src[i].  Strides in bytes. */
int sh_permute_rows(const void *src, int64_t src_ld_bytes, const int64_t *perm, int64_t n,
                    int64_t row_bytes, void *dst, int64_t dst_ld_bytes, int scatter,
                    sh_stream_t stream);

/* ---- fuzzy edge weights ---------------------------------------------------------
 * umap smooth_knn_dist + compute_membership_strengths per `chunk_size` row chunk,
 * then zero weights -> the global minimum rule of scarf.  dist FP32 [n x k],
 * idx i64 [n x k]; outputs FP32 sigma/rho [n], weights [n * k] row-major.
 * Replaces: knn_utils.py:89-159 (smoothen_dists). */
int sh_smooth_knn(const float *dist, const int64_t *idx, int64_t n, int32_t k,
                  int64_t chunk_size, float local_connectivity, float bandwidth,
                  int32_t n_iter, float *sigma_out, float *rho_out, float *w_out,
                  void *workspace, size_t workspace_bytes, sh_stream_t stream);
size_t sh_smooth_knn_workspace(int64_t n, int64_t chunk_size);

/* ---- mapping scores ---------------------------------------------------------------
 * For each target group g: ms[ref] = sum over the group's (query, neighbour) pairs of
 * 1 / (log1p(d) + 1) (or fixed_weight), scaled by multiplier / (n_g * k), then log1p.
 * Accumulated in 2^-40 fixed point with integer atomics: order-independent and
 * deterministic.  Target tables too large for one 64-bit fixed-point sum (more than
 * ~1.6e7 / k rows at weight <= 1) are taken in batches whose sums are added in FP64 in
 * batch order.  Distances are squared L2 (>= 0).  groups NULL: one group.
 * out FP64 [n_groups x n_ref].
 * Replaces: mapping_datastore.py:211-285 (get_mapping_score). */
int sh_mapping_score(const int64_t *idx, const float *dist, int64_t nq, int32_t k,
                     const int32_t *groups, int32_t n_groups, int64_t n_ref,
                     int weighted, double fixed_weight, double multiplier,
                     int log_transform, double *out, void *workspace,
                     size_t workspace_bytes, sh_stream_t stream);
size_t sh_mapping_score_workspace(int64_t n_ref, int32_t n_groups);
/* sh_mapping_score without a workspace argument: the accumulators live in a file-scope
 * device array of the library (64 MiB + 16 KiB per GPU, held until the process ends), so
 * the call allocates nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  Limits:
 * n_ref * n_groups <= 8,388,608 and n_groups <= 4096 (larger calls return SH_EINVAL;
 * sh_mapping_score_static_max(n_groups) gives the largest n_ref, 0 for a bad n_groups).
 * All calls on one GPU share the array: calls that overlap on one GPU (other streams or
 * host threads) race and must be serialised by the caller. */
int sh_mapping_score_static(const int64_t *idx, const float *dist, int64_t nq, int32_t k,
                            const int32_t *groups, int32_t n_groups, int64_t n_ref,
                            int weighted, double fixed_weight, double multiplier,
                            int log_transform, double *out, sh_stream_t stream);
size_t sh_mapping_score_static_max(int32_t n_groups);

/* ---- label transfer ----------------------------------------------------------------
 * Per target row of a mapping table (idx i64 [n x k], dist FP64 [n x k], as the
 * projections store holds them): weights w = 1 - d / max(d over the row); the first k - 1
 * neighbours vote with w for their reference class groups[idx] (i32 codes, [n_ref]); the
 * row's class is the single class whose vote divided by the sum of the k - 1 weights
 * (numpy's summation order) exceeds `threshold`, else -1.  k <= 129.
 * Replaces: mapping_datastore.py:287-366 (get_target_classes). */
int sh_target_classes(const int64_t *idx, const double *dist, int64_t n, int32_t k,
                      const int32_t *groups, int64_t n_ref, double threshold, int32_t *out,
                      sh_stream_t stream);

/* ---- MiniBatchKMeans seed partitions --------------------------------------------------
 * scarf's AnnStream._fit_kmeans (ann.py:328-346) runs scikit-learn 1.7.2's
 * MiniBatchKMeans: one partial_fit per block of cells (sklearn cluster/_kmeans.py:2203-2303),
 * then predict on every block (_kmeans.py:1066-1098).  The host keeps the random state and
 * makes the draws in scikit-learn's order; these entry points do the arithmetic in FP64.
 * Rows X are FP32 or FP64 (x_dtype SH_F32 / SH_F64), row-major with leading dimension ldx;
 * centers are FP64 [n_clusters x d], contiguous.
 *
 * sh_kmeans_pp: k-means++ seeding of n_clusters centers among the n points
 * X[rows[i]] (or X[i] if rows is NULL), replacing _kmeans.py:174-272 (_kmeans_plusplus).
 * `first` is the first center (the host's random_state.choice), `uniforms` the
 * (n_clusters - 1) x n_trials draws of random_state.uniform in order.  Writes the chosen
 * point indices (0..n-1) to centers_out.  d <= 256, n_trials <= 16; one workgroup. */
size_t sh_kmeans_pp_workspace(int64_t n, int32_t n_trials);
int sh_kmeans_pp(const void *X, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                 const int64_t *rows, int32_t n_clusters, int32_t n_trials, int64_t first,
                 const double *uniforms, int64_t *centers_out, void *workspace,
                 size_t workspace_bytes, sh_stream_t stream);
/* labels[i] = argmin_j ||c_j||^2 - 2 x_i.c_j, the lowest j on ties; center_norms [n_clusters]
 * receives ||c_j||^2.  Replaces _k_means_lloyd.pyx:172-215 as called by _labels_inertia
 * (_kmeans.py:755-822) in partial_fit and predict.  With a workspace of
 * sh_kmeans_assign_workspace bytes (0 for large n) few rows are spread over more
 * workgroups; the labels do not depend on it. */
size_t sh_kmeans_assign_workspace(int64_t n, int32_t n_clusters);
int sh_kmeans_assign(const void *X, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                     const double *centers, int32_t n_clusters, double *center_norms,
                     int32_t *labels, void *workspace, size_t workspace_bytes,
                     sh_stream_t stream);
/* In-place minibatch update of every center with its members among the n rows (unit
 * weights): c = (c * w + sum of members in row order) * (1 / (w + count)), w += count,
 * unchanged when count is 0; counts[j] receives the member count.  Same roundings as
 * _k_means_minibatch.pyx:59-108 (update_center_dense).  d <= 1024. */
int sh_kmeans_update(const void *X, int x_dtype, int64_t n, int32_t d, int64_t ldx,
                     const int32_t *labels, double *centers, int32_t n_clusters,
                     double *weight_sums, int32_t *counts, sh_stream_t stream);
/* centers[dst_rows[r]] = X[src_rows[r]] for r < m (dst_rows NULL: r).  Places the seeds
 * (_kmeans.py:1003-1020) and the random reassignments (_kmeans.py:1642-1673). */
int sh_kmeans_gather(const void *X, int x_dtype, int32_t d, int64_t ldx,
                     const int64_t *src_rows, const int32_t *dst_rows, int64_t m,
                     double *centers, sh_stream_t stream);

/* ---- cell graph as a sparse matrix ----------------------------------------------------
 * The stored kNN graph (edges i64 [nnz x 2] as (row, col), weights f64 [nnz], read in
 * groups of k entries, the first use_k of each group kept) as a canonical CSR g: columns
 * sorted, duplicate entries summed in input order, explicit zeros kept.  symmetric != 0:
 * t = g + g^T - g o g^T with scipy's sparse arithmetic (each operation drops the zeros it
 * produces); upper_only then keeps col >= row.  Writes indptr i64 [n + 1] (indptr[n] is the
 * entry count), indices i32 and data f64 (capacity >= kept edges, twice that when
 * symmetric) and *n_bad (device u32): the number of edges with a row or column outside
 * [0, n), which are left out.  Replaces graph_datastore.py:474-511 (_store_to_sparse) and
 * 1022-1075 (load_graph). */
size_t sh_graph_csr_workspace(int64_t n, int64_t nnz, int32_t k, int32_t use_k);
int sh_graph_csr(const int64_t *edges, const double *weights, int64_t nnz, int32_t k,
                 int32_t use_k, int64_t n, int symmetric, int upper_only, int64_t *indptr,
                 int32_t *indices, double *data, int64_t capacity, unsigned int *n_bad,
                 void *workspace, size_t workspace_bytes, sh_stream_t stream);

/* ---- CORAL alignment -------------------------------------------------------------------
 * C = (G - n m m^T) / (n - 1) + ridge I with m = colsum / n, +inf entries set to 0: the
 * covariance (ddof 1) of n rows from their Gram G (FP64, leading dimension ldg) and column
 * sums, as _cov_diaged (mapping_utils.py:15-19) with ridge 1. */
int sh_cov_from_gram(const double *G, int64_t ldg, const double *colsum, int64_t n, int32_t d,
                     double ridge, double *C, sh_stream_t stream);
/* C^{1/2} and C^{-1/2} of a symmetric positive definite C (FP64 [d x d]) by the coupled
 * Newton-Schulz iteration on FP64 MFMA GEMMs, replacing scipy's fractional_matrix_power
 * (mapping_utils.py:22-38).  Stops when ||I - Z Y||_F <= tol sqrt(d) or when the residual
 * stops falling below 1e-6, after at most max_iter iterations; *iters and *resid report the
 * count and the last residual.  Reads the residual back once per iteration (synchronous). */
size_t sh_spd_sqrt_workspace(int32_t d);
int sh_spd_sqrt(const double *C, int32_t d, double *sqrt_out, double *inv_sqrt_out,
                int32_t max_iter, double tol, int32_t *iters, double *resid, void *workspace,
                size_t workspace_bytes, sh_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* SCARF_HIP_H */
