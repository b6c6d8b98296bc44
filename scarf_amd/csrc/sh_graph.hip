// The cell graph as a sparse matrix: scarf's load_graph (graph_datastore.py:1022-1075) over
// _store_to_sparse (graph_datastore.py:474-511).  The stored edges [nnz x 2] (row, col) and
// weights are read in groups of k (one kNN row each), the first use_k of every group kept,
// and turned into a canonical CSR g (columns sorted, duplicates summed in input order,
// explicit zeros kept, as scipy's coo -> csr).  symmetric: t = g + g^T - g o g^T with
// scipy's sparse arithmetic (each of the three operations drops the zeros it produces),
// upper_only: t restricted to col >= row.
//
// Entries are sorted as 64-bit keys (row << cb | col, cb = bits of n) with hipCUB's stable
// radix sort; the transpose is the same sort on (col << cb | row).  Each output row of t is
// a merge of a row of g with the matching row of g^T, counted, scanned, then written.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "sh_common.h"

namespace sh {

__global__ void k_g_keys(const int64_t *__restrict__ edges, const double *__restrict__ w,
                         int64_t m, int32_t k, int32_t use_k, int64_t n, int cb,
                         uint64_t sentinel, uint64_t *__restrict__ keys,
                         double *__restrict__ vals, unsigned int *__restrict__ bad) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= m) return;
  const int64_t e = (o / use_k) * k + (o % use_k);
  const int64_t r = edges[2 * e], c = edges[2 * e + 1];
  if (r < 0 || r >= n || c < 0 || c >= n) {
    keys[o] = sentinel;
    vals[o] = 0.0;
    atomicAdd(bad, 1u);
    return;
  }
  keys[o] = ((uint64_t)r << cb) | (uint64_t)c;
  vals[o] = w[e];
}

__global__ void k_g_fill(uint64_t *__restrict__ keys, int64_t m, uint64_t sentinel,
                         int64_t *__restrict__ cnt_a, int64_t *__restrict__ cnt_b,
                         int64_t n1, unsigned int *__restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0 && bad) *bad = 0u;
  if (t < m) keys[t] = sentinel;
  if (t < n1) {
    cnt_a[t] = 0;
    cnt_b[t] = 0;
  }
}

__global__ void k_g_heads(const uint64_t *__restrict__ keys, int64_t m, uint64_t sentinel,
                          int32_t *__restrict__ head) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= m) return;
  const uint64_t key = keys[o];
  head[o] = key != sentinel && (o == 0 || keys[o - 1] != key) ? 1 : 0;
}

// one thread per run of equal keys: the run's values summed in input order (the radix
// sort is stable); g's row/col/value and the transposed key land at the run's position
__global__ void k_g_compact(const uint64_t *__restrict__ keys, const double *__restrict__ vals,
                            const int32_t *__restrict__ head, const int32_t *__restrict__ pos,
                            int64_t m, int cb, int32_t *__restrict__ g_row,
                            int32_t *__restrict__ g_col, double *__restrict__ g_val,
                            uint64_t *__restrict__ keyT, double *__restrict__ valT,
                            int64_t *__restrict__ rowcnt, int64_t *__restrict__ colcnt) {
#pragma clang fp contract(off)
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= m || !head[o]) return;
  const uint64_t key = keys[o];
  double x = vals[o];
  for (int64_t q = o + 1; q < m && keys[q] == key; ++q) x += vals[q];
  const uint64_t mask = (1ull << cb) - 1ull;
  const int32_t r = (int32_t)(key >> cb), c = (int32_t)(key & mask);
  const int64_t p = (int64_t)pos[o] - 1;
  g_row[p] = r;
  g_col[p] = c;
  g_val[p] = x;
  keyT[p] = ((uint64_t)c << cb) | (uint64_t)r;
  valT[p] = x;
  atomicAdd((unsigned long long *)&rowcnt[r], 1ull);
  atomicAdd((unsigned long long *)&colcnt[c], 1ull);
}

// t = g + g^T - g o g^T for one row, with scipy's zero dropping at each operation.
// write == false counts the row's entries; write == true stores them.
template <bool kWrite>
__global__ void k_g_sym(const int64_t *__restrict__ gp, const int32_t *__restrict__ g_col,
                        const double *__restrict__ g_val, const int64_t *__restrict__ tp,
                        const uint64_t *__restrict__ t_key, const double *__restrict__ t_val,
                        int64_t n, int cb, int upper_only, int64_t *__restrict__ cnt,
                        const int64_t *__restrict__ op, int32_t *__restrict__ out_col,
                        double *__restrict__ out_val) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (!kWrite && i == 0) cnt[n] = 0;  // the scan's last slot: the total
  if (i >= n) return;
  const uint64_t mask = (1ull << cb) - 1ull;
  int64_t a = gp[i], ae = gp[i + 1], b = tp[i], be = tp[i + 1];
  int64_t w = kWrite ? op[i] : 0;
  int64_t count = 0;
  while (a < ae || b < be) {
    const int64_t ca = a < ae ? (int64_t)g_col[a] : INT64_MAX;
    const int64_t cbv = b < be ? (int64_t)(t_key[b] & mask) : INT64_MAX;
    int64_t col;
    bool has_a = false, has_b = false;
    double va = 0.0, vb = 0.0;
    if (ca <= cbv) {
      col = ca;
      has_a = true;
      va = g_val[a++];
    } else {
      col = cbv;
    }
    if (cbv == col) {
      has_b = true;
      vb = t_val[b++];
    }
    // g + g^T: a + b, or the one present; dropped when zero
    double s = has_a && has_b ? va + vb : (has_a ? va : vb);
    const bool s_in = s != 0.0;
    // g o g^T: only where both are present; dropped when zero
    const double pr = has_a && has_b ? va * vb : 0.0;
    const bool p_in = has_a && has_b && pr != 0.0;
    if (!s_in && !p_in) continue;
    const double t = (s_in ? s : 0.0) - (p_in ? pr : 0.0);
    if (t == 0.0) continue;
    if (upper_only && col < i) continue;
    if (kWrite) {
      out_col[w] = (int32_t)col;
      out_val[w] = t;
      ++w;
    } else {
      ++count;
    }
  }
  if (!kWrite) cnt[i] = count;
}

__global__ void k_g_copy_csr(const int64_t *__restrict__ gp, const int32_t *__restrict__ g_col,
                             const double *__restrict__ g_val, int64_t n, int64_t m,
                             int64_t *__restrict__ indptr, int32_t *__restrict__ indices,
                             double *__restrict__ data) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= n) indptr[t] = gp[t];
  if (t < m && t < gp[n]) {
    indices[t] = g_col[t];
    data[t] = g_val[t];
  }
}

inline int key_bits(int64_t n) {
  int b = 1;
  while (b < 31 && ((int64_t)1 << b) <= n) ++b;  // 2^b > n, so the sentinel is never a key
  return b;
}

struct GraphWs {
  uint64_t *keys, *keys_s;
  double *vals, *vals_s;
  int32_t *head, *pos, *g_row, *g_col;
  double *g_val;
  int64_t *rowcnt, *colcnt, *gp, *tp, *cnt, *op;
  unsigned int *bad;
  void *tmp;
  size_t tmp_bytes;
};

inline size_t graph_tmp_bytes(int64_t m, int64_t n) {
  size_t a = 0, b = 0, c = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const uint64_t *)nullptr,
                                     (uint64_t *)nullptr, (const double *)nullptr,
                                     (double *)nullptr, (int)m, 0, 62);
  (void)hipcub::DeviceScan::InclusiveSum(nullptr, b, (const int32_t *)nullptr, (int32_t *)nullptr,
                                   (int)m);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, c, (const int64_t *)nullptr, (int64_t *)nullptr,
                                   (int)(n + 1));
  return std::max(a, std::max(b, c));
}

// carves the workspace; returns its size (p == nullptr: size only)
inline size_t graph_ws(unsigned char *p, int64_t m, int64_t n, GraphWs *w) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    void *q = p ? (void *)(p + off) : nullptr;
    off += align_up(bytes, 256);
    return q;
  };
  const size_t mm = (size_t)std::max<int64_t>(m, 1), n1 = (size_t)(n + 1);
  w->keys = (uint64_t *)take(mm * 8);
  w->keys_s = (uint64_t *)take(mm * 8);
  w->vals = (double *)take(mm * 8);
  w->vals_s = (double *)take(mm * 8);
  w->head = (int32_t *)take(mm * 4);
  w->pos = (int32_t *)take(mm * 4);
  w->g_row = (int32_t *)take(mm * 4);
  w->g_col = (int32_t *)take(mm * 4);
  w->g_val = (double *)take(mm * 8);
  w->rowcnt = (int64_t *)take(n1 * 8);
  w->colcnt = (int64_t *)take(n1 * 8);
  w->gp = (int64_t *)take(n1 * 8);
  w->tp = (int64_t *)take(n1 * 8);
  w->cnt = (int64_t *)take(n1 * 8);
  w->op = (int64_t *)take(n1 * 8);
  w->bad = (unsigned int *)take(256);
  w->tmp_bytes = graph_tmp_bytes(mm, n);
  w->tmp = take(w->tmp_bytes);
  return off;
}

}  // namespace sh

using namespace sh;

extern "C" size_t sh_graph_csr_workspace(int64_t n, int64_t nnz, int32_t k, int32_t use_k) {
  if (n < 1 || k < 1 || use_k < 1 || use_k > k || nnz < 0 || nnz % k) return 0;
  GraphWs w;
  return graph_ws(nullptr, nnz / k * use_k, n, &w);
}

extern "C" int sh_graph_csr(const int64_t *edges, const double *weights, int64_t nnz,
                            int32_t k, int32_t use_k, int64_t n, int symmetric, int upper_only,
                            int64_t *indptr, int32_t *indices, double *data, int64_t capacity,
                            unsigned int *n_bad, void *ws, size_t ws_bytes,
                            sh_stream_t stream) {
  SH_ARG(n >= 1 && n < 0x7fffffff, "sh_graph_csr: need 1 <= n < 2^31");
  SH_ARG(k >= 1 && use_k >= 1 && use_k <= k, "sh_graph_csr: need 1 <= use_k <= k");
  SH_ARG(nnz >= 0 && nnz % k == 0, "sh_graph_csr: nnz must be a multiple of k");
  const int64_t m = nnz / k * use_k;
  SH_ARG(m < 0x7fffffff, "sh_graph_csr: too many edges");
  SH_ARG(capacity >= (symmetric ? 2 * m : m), "sh_graph_csr: output capacity below %s",
         symmetric ? "2 * kept edges" : "kept edges");
  SH_ARG(indptr && n_bad && ws && (m == 0 || (edges && weights && indices && data)),
         "sh_graph_csr: NULL pointer");
  GraphWs w;
  const size_t need = graph_ws(nullptr, m, n, &w);
  if (ws_bytes < need) {
    set_error("sh_graph_csr: workspace %zu bytes < %zu", ws_bytes, need);
    return SH_EWORKSPACE;
  }
  graph_ws((unsigned char *)ws, m, n, &w);
  hipStream_t s = (hipStream_t)stream;
  const int cb = key_bits(n);
  const uint64_t sentinel = (1ull << (2 * cb)) - 1ull;
  const int64_t mm = std::max<int64_t>(m, 1);
  hipLaunchKernelGGL(k_g_fill, dim3((uint32_t)cdiv(std::max(mm, n + 1), 256)), dim3(256), 0, s,
                     w.keys_s, (int64_t)0, sentinel, w.rowcnt, w.colcnt, n + 1, w.bad);
  SH_LAUNCH_CHECK("k_g_fill");
  if (m > 0) {
    hipLaunchKernelGGL(k_g_keys, dim3((uint32_t)cdiv(m, 256)), dim3(256), 0, s, edges, weights,
                       m, k, use_k, n, cb, sentinel, w.keys, w.vals, w.bad);
    SH_LAUNCH_CHECK("k_g_keys");
    size_t tb = w.tmp_bytes;
    SH_HIP(hipcub::DeviceRadixSort::SortPairs(w.tmp, tb, w.keys, w.keys_s, w.vals, w.vals_s,
                                              (int)m, 0, 2 * cb, s));
    hipLaunchKernelGGL(k_g_heads, dim3((uint32_t)cdiv(m, 256)), dim3(256), 0, s, w.keys_s, m,
                       sentinel, w.head);
    SH_LAUNCH_CHECK("k_g_heads");
    tb = w.tmp_bytes;
    SH_HIP(hipcub::DeviceScan::InclusiveSum(w.tmp, tb, w.head, w.pos, (int)m, s));
    // transposed keys start as sentinels: the slots past the unique count sort last
    hipLaunchKernelGGL(k_g_fill, dim3((uint32_t)cdiv(m, 256)), dim3(256), 0, s, w.keys, m,
                       sentinel, (int64_t *)nullptr, (int64_t *)nullptr, (int64_t)0,
                       (unsigned int *)nullptr);
    SH_LAUNCH_CHECK("k_g_fill");
    hipLaunchKernelGGL(k_g_compact, dim3((uint32_t)cdiv(m, 256)), dim3(256), 0, s, w.keys_s,
                       w.vals_s, w.head, w.pos, m, cb, w.g_row, w.g_col, w.g_val, w.keys,
                       w.vals, w.rowcnt, w.colcnt);
    SH_LAUNCH_CHECK("k_g_compact");
  }
  size_t tb = w.tmp_bytes;
  SH_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.rowcnt, w.gp, (int)(n + 1), s));
  if (!symmetric) {
    hipLaunchKernelGGL(k_g_copy_csr, dim3((uint32_t)cdiv(std::max(mm, n + 1), 256)), dim3(256),
                       0, s, w.gp, w.g_col, w.g_val, n, m, indptr, indices, data);
    SH_LAUNCH_CHECK("k_g_copy_csr");
  } else {
    tb = w.tmp_bytes;
    SH_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.colcnt, w.tp, (int)(n + 1), s));
    if (m > 0) {
      tb = w.tmp_bytes;
      SH_HIP(hipcub::DeviceRadixSort::SortPairs(w.tmp, tb, w.keys, w.keys_s, w.vals, w.vals_s,
                                                (int)m, 0, 2 * cb, s));
    }
    const dim3 grid((uint32_t)cdiv(n, 256));
    hipLaunchKernelGGL(k_g_sym<false>, grid, dim3(256), 0, s, w.gp, w.g_col, w.g_val, w.tp,
                       w.keys_s, w.vals_s, n, cb, upper_only, w.cnt, (const int64_t *)nullptr,
                       (int32_t *)nullptr, (double *)nullptr);
    SH_LAUNCH_CHECK("k_g_sym<count>");
    tb = w.tmp_bytes;
    SH_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.cnt, indptr, (int)(n + 1), s));
    hipLaunchKernelGGL(k_g_sym<true>, grid, dim3(256), 0, s, w.gp, w.g_col, w.g_val, w.tp,
                       w.keys_s, w.vals_s, n, cb, upper_only, (int64_t *)nullptr, indptr,
                       indices, data);
    SH_LAUNCH_CHECK("k_g_sym<write>");
  }
  SH_HIP(hipMemcpyAsync(n_bad, w.bad, sizeof(unsigned int), hipMemcpyDeviceToDevice, s));
  return SH_OK;
}
