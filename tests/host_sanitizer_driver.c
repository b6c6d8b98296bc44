/* Host-side checks of libscarfhip's C ABI under AddressSanitizer and
 * UndefinedBehaviorSanitizer (built host-only by tests/test_host_sanitizers.py; no GPU is
 * touched).  Every *_workspace size function runs over a grid of sizes, and every entry
 * point is called with arguments its validation must reject (NULL pointers with
 * non-zero sizes, negative or inconsistent sizes): each call has to return an error code
 * with a message, and the sanitizers report any out-of-bounds access, leak of the error
 * buffer or arithmetic overflow on the way.  Prints "ok <n>" with the number of checked
 * calls, exits non-zero on the first failure. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "scarf_hip.h"

static int n_checked = 0;

static void expect_error(int rc, const char *what) {
  const char *msg = sh_last_error();
  if (rc >= 0 || msg == NULL || msg[0] == '\0') {
    fprintf(stderr, "FAIL %s: rc=%d msg=\"%s\"\n", what, rc, msg ? msg : "(null)");
    exit(1);
  }
  ++n_checked;
}

#define BAD(call) expect_error((call), #call)

static volatile size_t sink = 0;

int main(void) {
  if (sh_version() <= 0) {
    fprintf(stderr, "FAIL sh_version\n");
    return 1;
  }
  /* ---- workspace sizes over a grid (including zero and negative sizes) ---- */
  const int64_t ns[] = {-1, 0, 1, 255, 256, 257, 32768, 1000000, 4000000, 100000000};
  const int32_t ds[] = {-1, 0, 1, 13, 50, 128, 2000, 3000};
  for (size_t i = 0; i < sizeof(ns) / sizeof(ns[0]); ++i) {
    const int64_t n = ns[i];
    for (size_t j = 0; j < sizeof(ds) / sizeof(ds[0]); ++j) {
      const int32_t d = ds[j];
      sink += sh_norm_stats_workspace(n, 50000, d);
      sink += sh_norm_stats_workspace(n, -1, d);
      sink += sh_gram_workspace(n, d, 1);
      sink += sh_gram_workspace(n, d, 0);
      sink += sh_sparse_gram_workspace(n, n > 0 ? n * 21 : 0, d);
      sink += sh_lanczos_workspace(d, 192);
      sink += sh_block_gram_workspace(d, 94, 94);
      sink += sh_symm_block_workspace(d, 96, 96);
      sink += sh_symm_block_workspace(d, 16, 192);
      sink += sh_knn_screen_workspace(n, 64, 1);
      sink += sh_knn_screen_workspace(n, 256, 4);
      sink += sh_locality_order_workspace(n);
      sink += sh_smooth_knn_workspace(n, 1000);
      sink += sh_mapping_score_workspace(n, d);
      sink += sh_mapping_score_static_max(d);
      sink += sh_kmeans_pp_workspace(n, 16);
      sink += sh_kmeans_assign_workspace(n, 1000);
      sink += sh_graph_csr_workspace(n, n * 21, 21, 11);
      sink += sh_spd_sqrt_workspace(d);
      sink += (size_t)sh_knn_pack_k(SH_BF16, d) + (size_t)sh_knn_pack_row_bytes(SH_FP8, d);
      n_checked += 21;
    }
  }
  if (sh_knn_pack_k(99, 50) >= 0 || sh_knn_pack_row_bytes(99, 50) >= 0) {
    fprintf(stderr, "FAIL unknown dtype accepted\n");
    return 1;
  }

  /* ---- every entry point rejects what it must ---- */
  void *ws = NULL;
  const size_t wsb = 0;
  sh_stream_t st = NULL;
  BAD(sh_norm_stats(NULL, NULL, NULL, 100, 10, NULL, 50, 50, 1e4, 1, NULL, NULL, NULL, NULL,
                    ws, wsb, st));
  BAD(sh_norm_stats(NULL, NULL, NULL, -5, 10, NULL, 50, 50, 1e4, 1, NULL, NULL, NULL, NULL,
                    ws, wsb, st));
  BAD(sh_stats_merge(NULL, NULL, NULL, 4, 50, NULL, NULL, NULL, st));
  BAD(sh_densify(NULL, NULL, NULL, NULL, 10, NULL, 50, 50, 1e4, 1, NULL, NULL, NULL, 0.0, 1,
                 NULL, 127, st));
  BAD(sh_gram(NULL, 1, 100, 50, 100, 1, 0, NULL, NULL, ws, wsb, st));
  BAD(sh_gram(NULL, 1, -1, 50, 128, 1, 0, NULL, NULL, ws, wsb, st));
  BAD(sh_sparse_gram(NULL, NULL, NULL, NULL, 10, 100, NULL, 50, 50, 1e4, 1, NULL, NULL, NULL,
                     0.0, 0, NULL, NULL, ws, wsb, st));
  BAD(sh_sparse_gram((const uint32_t *)8, (const int32_t *)8, (const int64_t *)8, NULL, 10, 100,
                     NULL, 9000, 9000, 1e4, 1, (const double *)8, (const double *)8,
                     (const double *)8, 0.0, 0, (double *)8, (double *)8, ws, wsb, st));
  BAD(sh_sparse_gram((const uint32_t *)8, (const int32_t *)8, (const int64_t *)8, NULL, 10, 100,
                     NULL, 50, 50, 1e4, 1, (const double *)8, (const double *)8,
                     (const double *)8, 0.0, 0, (double *)8, (double *)8, ws, wsb, st));
  BAD(sh_merge_matrix(NULL, 50, 10.0, NULL, NULL, NULL, 0, NULL, NULL, st));
  BAD(sh_lanczos_steps(NULL, 50, NULL, 64, 0, 2, NULL, NULL, NULL, ws, wsb, st));
  BAD(sh_dgemm(0, -1, 4, 4, 1.0, NULL, 4, NULL, 4, 0.0, NULL, 4, st));
  BAD(sh_dgemm(0, 4, 4, 4, 1.0, NULL, 4, NULL, 4, 0.0, NULL, 4, st));
  BAD(sh_block_gram(NULL, 96, 94, NULL, 96, 94, 2000, NULL, ws, wsb, st));
  BAD(sh_symm_block(NULL, 2000, NULL, NULL, 96, 96, 1.0, 0.0, 0.0, NULL, ws, wsb, st));
  BAD(sh_symm_block((const double *)8, 2000, (const double *)16, NULL, 40, 48, 1.0, 0.0, 1.0,
                    (double *)32, (void *)64, 1 << 20, st));
  BAD(sh_block_residual(NULL, NULL, 2000, 94, 96, NULL, NULL, st));
  BAD(sh_dense_z(NULL, SH_F64, 10, 50, 50, NULL, NULL, 0.0, NULL, 50, st));
  BAD(sh_dense_project(NULL, SH_F32, 10, 50, 50, NULL, NULL, 0.0, NULL, 200, NULL, 200, st));
  BAD(sh_project(NULL, NULL, NULL, 10, NULL, 50, 50, 1e4, 1, NULL, NULL, NULL, 0.f, NULL, 52,
                 50, NULL, NULL, 50, st));
  BAD(sh_project(NULL, NULL, NULL, 10, NULL, 50, 50, 1e4, 1, NULL, NULL, NULL, 0.f, NULL, 10,
                 50, NULL, NULL, 50, st));
  BAD(sh_knn_pack(NULL, 10, 50, 50, SH_BF16, 0, NULL, NULL, st));
  BAD(sh_knn_pack(NULL, 10, 50, 50, 99, 0, NULL, NULL, st));
  BAD(sh_knn_screen(NULL, 10, NULL, 10, 50, SH_BF16, NULL, NULL, 64, 1, NULL, NULL, NULL, ws,
                    wsb, st));
  BAD(sh_knn_screen(NULL, 10, NULL, 10, 50, SH_BF16, NULL, NULL, 64, 0, NULL, NULL, NULL, ws,
                    wsb, st));
  BAD(sh_knn_rerank(NULL, 10, 50, NULL, 10, 50, 50, NULL, NULL, 64, 1, 11, -1, NULL, -1.0, 0.f,
                    0.0, NULL, NULL, NULL, st));
  BAD(sh_knn_rerank(NULL, 10, 50, NULL, 10, 50, 50, NULL, NULL, 8, 1, 11, -1, NULL, -1.0, 0.f,
                    0.0, NULL, NULL, NULL, st));
  BAD(sh_knn_exact(NULL, 10, 50, NULL, 10, NULL, 10, 50, 50, 11, -1, NULL, NULL, NULL, NULL,
                   NULL, st));
  BAD(sh_knn_exact(NULL, 10, 50, NULL, 10, NULL, 10, 50, 50, 100, -1, NULL, NULL, NULL, NULL,
                   NULL, st));
  BAD(sh_fix_knn_query(NULL, NULL, NULL, 10, 12, NULL, NULL, NULL, st));
  BAD(sh_locality_order(NULL, 10, 50, 50, NULL, ws, wsb, st));
  BAD(sh_permute_rows(NULL, 200, NULL, 10, 200, NULL, 200, 0, st));
  BAD(sh_smooth_knn(NULL, NULL, 10, 11, 5, 1.f, 1.5f, 64, NULL, NULL, NULL, ws, wsb, st));
  BAD(sh_mapping_score(NULL, NULL, 10, 11, NULL, 1, 100, 0, 0.1, 1000.0, 1, NULL, ws, wsb, st));
  BAD(sh_mapping_score_static(NULL, NULL, 10, 11, NULL, 1, 100, 0, 0.1, 1000.0, 1, NULL, st));
  BAD(sh_mapping_score_static((const int64_t *)8, (const float *)8, 10, 11, NULL, 4,
                              (int64_t)1 << 40, 0, 0.1, 1000.0, 1, (double *)8, st));
  BAD(sh_target_classes(NULL, NULL, 10, 11, NULL, 100, 0.5, NULL, st));
  BAD(sh_kmeans_pp(NULL, SH_F32, 100, 50, 50, NULL, 10, 3, 0, NULL, NULL, ws, wsb, st));
  BAD(sh_kmeans_assign(NULL, SH_F32, 100, 50, 50, NULL, 10, NULL, NULL, ws, wsb, st));
  BAD(sh_kmeans_update(NULL, SH_F32, 100, 50, 50, NULL, NULL, 10, NULL, NULL, st));
  BAD(sh_kmeans_gather(NULL, SH_F32, 50, 50, NULL, NULL, 4, NULL, st));
  BAD(sh_graph_csr(NULL, NULL, 100, 11, 11, 10, 1, 0, NULL, NULL, NULL, 100, NULL, ws, wsb,
                   st));
  BAD(sh_cov_from_gram(NULL, 50, NULL, 100, 50, 1.0, NULL, st));
  BAD(sh_spd_sqrt(NULL, 50, NULL, NULL, 10, 1e-12, NULL, NULL, ws, wsb, st));

  /* the last message fits its buffer */
  if (strlen(sh_last_error()) >= 1024) {
    fprintf(stderr, "FAIL error message longer than its buffer\n");
    return 1;
  }
  printf("ok %d (sink %zu)\n", n_checked, (size_t)(sink & 1));
  return 0;
}
