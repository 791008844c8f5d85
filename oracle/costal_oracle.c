/*
 * costal_oracle.c -- CPU restatement of CosTaL's exact cosine kNN and edge weights.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in costal_amd/:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and only to
 * compare against or to time as the CPU baseline.  Nothing in the product path calls it.
 *
 * What it restates (reference = ike12th/CosTaL 2.1.0):
 *   oracle_row_norms64        crates/costal-core/src/knn.rs:289-296, 315-323  (dot64, row_norms64)
 *   cos64                     crates/costal-core/src/knn.rs:300-312
 *   validate                  crates/costal-core/src/knn.rs:325-342
 *   oracle_knn_bruteforce     crates/costal-core/src/knn.rs:884-924          (the definition)
 *   oracle_knn_blocked        crates/costal-core/src/knn.rs:403-512, 545-569, 603-660, 688-875
 *                             (plain f32 coarse blocks + heaps + f64 rescoring + epsilon
 *                             certificate + gather_row fallback + unpermute)
 *   oracle_tani_from_cos      crates/costal-core/src/weight.rs:28-37
 *   oracle_reweight           crates/costal-core/src/weight.rs:41-93 (tanimoto, cosine, jaccard)
 *   oracle_fuzzy_union        crates/costal-core/src/graph.rs:24-68
 *   oracle_csr_from_edges     crates/costal-core/src/leiden.rs:52-100
 *   oracle_heap_keep          crates/costal-core/src/knn.rs:791-875 (BlockHeaps, for its own test)
 *
 * How each is pinned (tests/test_oracle.py):
 *   - closed forms the reference's tests hold: duplicate prototypes -> row 0 = [10,20,...,100]
 *     (knn.rs:1001-1017), zero rows -> cos 0 (knn.rs:1030-1031), heap keeps {0,3,9}
 *     (knn.rs:1149-1159), tani_from_cos(1,2,1)=2/3 etc. (weight.rs:99-109), Jaccard and fuzzy-union
 *     vectors (weight.rs:111-121, graph.rs:74-95), the error cases (knn.rs:1046-1067);
 *   - an independent numpy float64 matmul oracle at atol 1e-12 plus identical neighbor sets on
 *     clear-gap rows (tests/test_knn.py:46-61);
 *   - bitwise agreement of the sequential f64 sums with numpy's sequential np.cumsum on small inputs;
 *   - brute force == blocked form for many tile shapes, layouts and thread counts
 *     (knn.rs:960-999, tests/test_knn.py:38-94) -- two independent code paths.
 *
 * Numerics: this translation unit must be compiled with -ffp-contract=off so that
 * `a*a + b*b - d` and friends round exactly as written (Rust never contracts).  The f32 tile
 * product lives in oracle_gemm.c, which may contract: it feeds only the coarse stage, whose error
 * the certificate bounds.
 */
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "costal_oracle.h"

/* ------------------------------------------------------------------ errors */

static _Thread_local char g_err[256];

const char* oracle_last_error(void) { return g_err; }

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ------------------------------------------------------------------ f64 primitives */

/* knn.rs:289-296: f64 sum in index order; f32*f32 is exact in f64. */
double oracle_dot64(const float* a, const float* b, uint64_t d) {
  double s = 0.0;
  for (uint64_t t = 0; t < d; ++t) s += (double)a[t] * (double)b[t];
  return s;
}

/* knn.rs:300-312 */
static double cos64(double dot, double ni, double nj) {
  if (ni == 0.0 || nj == 0.0) return 0.0;
  double c = dot / (ni * nj);
  return c == 0.0 ? 0.0 : c;
}

/* weight.rs:28-37 */
double oracle_tani_from_cos(double a, double b, double cos) {
  double d = a * b * cos;
  if (d == 0.0) return 0.0;
  return d / (a * a + b * b - d);
}

/* knn.rs:315-323 */
int oracle_row_norms64(const float* x, uint64_t n, uint64_t d, double* norms, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const float* r = x + (uint64_t)i * d;
    norms[i] = sqrt(oracle_dot64(r, r, d));
  }
  return 0;
}

/* knn.rs:325-342.  Length is implicit in (n, d) on the C side. */
static int validate(const float* x, uint64_t n, uint64_t d, uint64_t k) {
  if (n > 0xffffffffull) return fail(ORACLE_E_TOO_MANY_ROWS, "n=%llu exceeds the u32 index range", (unsigned long long)n);
  if (k == 0 || k >= n) return fail(ORACLE_E_BAD_K, "invalid k=%llu: need 1 <= k <= n-1 (n=%llu)", (unsigned long long)k, (unsigned long long)n);
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t t = 0; t < d; ++t)
      if (!isfinite(x[i * d + t])) return fail(ORACLE_E_NON_FINITE, "row %llu contains NaN or inf", (unsigned long long)i);
  return 0;
}

/* ------------------------------------------------------------------ candidate ordering */

typedef struct {
  double score; /* selection score (cosine) */
  double cos;
  uint32_t j; /* original index */
} cand_t;

/* knn.rs:589-591: (score desc by total order, original index asc).  Scores here are never NaN
 * and -0.0 was normalized away by cos64, so '<' agrees with total_cmp. */
static int cand_cmp(const void* pa, const void* pb) {
  const cand_t* a = (const cand_t*)pa;
  const cand_t* b = (const cand_t*)pb;
  if (a->score > b->score) return -1;
  if (a->score < b->score) return 1;
  return (a->j < b->j) ? -1 : (a->j > b->j);
}

/* a ranks after b */
static inline int cand_after(const cand_t* a, const cand_t* b) { return cand_cmp(a, b) > 0; }

/* ------------------------------------------------------------------ brute force (the definition) */

/* knn.rs:884-924 with Select::Cosine: score every j != i by f64 cosine, keep the best k by
 * (score desc, j asc).  A bounded max-heap of the k best keeps memory O(k) per row. */
static void heap_sift_down(cand_t* h, uint64_t len, uint64_t pos) {
  /* root = worst (the one ranking last) */
  for (;;) {
    uint64_t l = 2 * pos + 1, w = pos;
    if (l < len && cand_after(&h[l], &h[w])) w = l;
    if (l + 1 < len && cand_after(&h[l + 1], &h[w])) w = l + 1;
    if (w == pos) return;
    cand_t t = h[pos];
    h[pos] = h[w];
    h[w] = t;
    pos = w;
  }
}

static void heap_sift_up(cand_t* h, uint64_t pos) {
  while (pos > 0) {
    uint64_t p = (pos - 1) / 2;
    if (!cand_after(&h[pos], &h[p])) return;
    cand_t t = h[pos];
    h[pos] = h[p];
    h[p] = t;
    pos = p;
  }
}

int oracle_knn_bruteforce(const float* x, uint64_t n, uint64_t d, uint64_t k, int select_tanimoto,
                          uint32_t* idx, double* cos_out, double* norms, int nthreads) {
  int rc = validate(x, n, d, k);
  if (rc) return rc;
  oracle_row_norms64(x, n, d, norms, nthreads);
  set_threads(nthreads);
  int oom = 0;
#pragma omp parallel
  {
    cand_t* h = (cand_t*)malloc(sizeof(cand_t) * k);
    if (!h) {
#pragma omp atomic write
      oom = 1;
    }
#pragma omp for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
      if (!h) continue;
      uint64_t i = (uint64_t)ii, len = 0;
      const float* xi = x + i * d;
      for (uint64_t j = 0; j < n; ++j) {
        if (j == i) continue;
        cand_t c;
        c.cos = cos64(oracle_dot64(xi, x + j * d, d), norms[i], norms[j]);
        c.score = select_tanimoto ? oracle_tani_from_cos(norms[i], norms[j], c.cos) : c.cos;
        c.j = (uint32_t)j;
        if (len < k) {
          h[len] = c;
          heap_sift_up(h, len);
          ++len;
        } else if (cand_after(&h[0], &c)) {
          h[0] = c;
          heap_sift_down(h, len, 0);
        }
      }
      qsort(h, len, sizeof(cand_t), cand_cmp);
      for (uint64_t t = 0; t < k; ++t) {
        idx[i * k + t] = h[t].j;
        cos_out[i * k + t] = h[t].cos;
      }
    }
    free(h);
  }
  if (oom) return fail(ORACLE_E_OOM, "out of memory");
  return 0;
}

/* ------------------------------------------------------------------ the blocked form */

/* BlockHeaps (knn.rs:791-875): per-row fixed-capacity min-heap; the root is the worst candidate:
 * lower score, or equal score and larger (working) index. */
static inline int worse(float sa, uint32_t ia, float sb, uint32_t ib) {
  return sa < sb || (sa == sb && ia > ib);
}

typedef struct {
  uint64_t cap;
  float* s;
  uint32_t* id;
  uint64_t* len;
} heaps_t;

static inline float heap_tau(const heaps_t* h, uint64_t r) {
  return h->len[r] < h->cap ? -INFINITY : h->s[r * h->cap];
}

static void heap_push(heaps_t* h, uint64_t r, float score, uint32_t id) {
  uint64_t cap = h->cap, len = h->len[r];
  float* s = h->s + r * cap;
  uint32_t* ids = h->id + r * cap;
  if (len < cap) {
    uint64_t pos = len;
    s[pos] = score;
    ids[pos] = id;
    while (pos > 0) {
      uint64_t parent = (pos - 1) / 2;
      if (!worse(s[pos], ids[pos], s[parent], ids[parent])) break;
      float ts = s[pos]; s[pos] = s[parent]; s[parent] = ts;
      uint32_t ti = ids[pos]; ids[pos] = ids[parent]; ids[parent] = ti;
      pos = parent;
    }
    h->len[r] = len + 1;
    return;
  }
  if (!worse(s[0], ids[0], score, id)) return;
  s[0] = score;
  ids[0] = id;
  uint64_t pos = 0;
  for (;;) {
    uint64_t left = 2 * pos + 1;
    if (left >= cap) break;
    uint64_t w = left;
    if (left + 1 < cap && worse(s[left + 1], ids[left + 1], s[left], ids[left])) w = left + 1;
    if (!worse(s[w], ids[w], s[pos], ids[pos])) break;
    float ts = s[pos]; s[pos] = s[w]; s[w] = ts;
    uint32_t ti = ids[pos]; ids[pos] = ids[w]; ids[w] = ti;
    pos = w;
  }
}

/* Test hook for the heap semantics (knn.rs:1149-1159): push (score, id) pairs into one heap of
 * capacity `cap`, write the kept ids (heap order) to `kept`, return how many. */
uint64_t oracle_heap_keep(const float* scores, const uint32_t* ids, uint64_t m, uint64_t cap, uint32_t* kept) {
  heaps_t h;
  uint64_t len = 0;
  h.cap = cap;
  h.s = (float*)malloc(sizeof(float) * cap);
  h.id = (uint32_t*)malloc(sizeof(uint32_t) * cap);
  h.len = &len;
  for (uint64_t t = 0; t < m; ++t) heap_push(&h, 0, scores[t], ids[t]);
  memcpy(kept, h.id, sizeof(uint32_t) * len);
  free(h.s);
  free(h.id);
  return len;
}

/* cache.rs:18-22 + knn.rs:551-569, with the tile budget passed in (speed only, never output). */
void oracle_block_sizes(uint64_t n, uint64_t row_block, uint64_t col_block, uint64_t budget_bytes,
                        int nthreads, uint64_t* rb_out, uint64_t* cb_out) {
  uint64_t budget = budget_bytes / sizeof(float);
  if (budget < 1) budget = 1;
  unsigned lg = 63u - (unsigned)__builtin_clzll(budget);
  uint64_t side = 1ull << (lg / 2);
  uint64_t cb = col_block ? col_block : (side < 256 ? 256 : side > 1024 ? 1024 : side);
  if (cb < 1) cb = 1;
  if (cb > n) cb = n;
  uint64_t rb;
  if (row_block) {
    rb = row_block;
  } else {
    uint64_t by_cache = budget / cb;
    uint64_t th = nthreads > 0 ? (uint64_t)nthreads : 1;
    uint64_t by_balance = (n + th * 8 - 1) / (th * 8);
    rb = by_cache < by_balance ? by_cache : by_balance;
    if (rb < 8) rb = 8;
    if (rb > 512) rb = 512;
  }
  if (rb < 1) rb = 1;
  if (rb > n) rb = n;
  *rb_out = rb;
  *cb_out = cb;
}

typedef struct {
  const float* x;
  const float* xs;
  const double* norms;
  const uint32_t* perm;
  uint64_t n, d, k, cap, cb;
  double eps;
} ctx_t;

/* knn.rs:746-753 */
static void push_rescored(const ctx_t* c, uint64_t i, uint64_t wj, cand_t* cands, uint64_t* nc) {
  uint32_t j = c->perm[wj];
  double dot = oracle_dot64(c->x + i * c->d, c->x + (uint64_t)j * c->d, c->d);
  double cs = cos64(dot, c->norms[i], c->norms[j]);
  cands[*nc].score = cs;
  cands[*nc].cos = cs;
  cands[*nc].j = j;
  ++*nc;
}

/* knn.rs:759-787: 16-wide chunk reject, rare heap insert; self excluded by working index. */
static void scan_row(const float* row, uint64_t cols, uint64_t c0, uint64_t self_col, heaps_t* h, uint64_t r) {
  const uint64_t LANES = 16;
  float th = heap_tau(h, r);
  for (uint64_t base = 0; base < cols; base += LANES) {
    uint64_t len = cols - base < LANES ? cols - base : LANES;
    int hit = 0;
    for (uint64_t t = 0; t < len; ++t) hit |= row[base + t] > th;
    if (!hit) continue;
    for (uint64_t t = 0; t < len; ++t) {
      float v = row[base + t];
      uint64_t col = c0 + base + t;
      if (v > th && col != self_col) {
        heap_push(h, r, v, (uint32_t)col);
        th = heap_tau(h, r);
      }
    }
  }
}

/* knn.rs:723-743: rescan every column for one row, gather f32 score >= min_sim, rescore. */
static void gather_row(const ctx_t* c, uint64_t w_row, double min_sim, float* tile, cand_t** cands,
                       uint64_t* nc, uint64_t* ccap) {
  uint64_t d = c->d, i = c->perm[w_row];
  const float* a = c->xs + w_row * d;
  *nc = 0;
  for (uint64_t c0 = 0; c0 < c->n; c0 += c->cb) {
    uint64_t cols = c->n - c0 < c->cb ? c->n - c0 : c->cb;
    oracle_tile_nt(a, 1, c->xs + c0 * d, cols, d, tile);
    for (uint64_t t = 0; t < cols; ++t) {
      uint64_t wj = c0 + t;
      if (wj != w_row && (double)tile[t] >= min_sim) {
        if (*nc == *ccap) {
          *ccap *= 2;
          *cands = (cand_t*)realloc(*cands, sizeof(cand_t) * *ccap);
        }
        push_rescored(c, i, wj, *cands, nc);
      }
    }
  }
}

/* The blocked exact kNN for working rows [row_begin, row_end) (the whole call when the range is
 * [0, n)).  With sort_by_norm the working order is norm-sorted and the output is unpermuted to the
 * original order, so a partial range is only meaningful with sort_by_norm = 0 (the CPU-baseline
 * sample).  stats[0..3] = fallback_rows, row_block, col_block, blocks_total. */
int oracle_knn_blocked(const float* x, uint64_t n, uint64_t d, uint64_t k, uint64_t margin,
                       uint64_t row_block, uint64_t col_block, uint64_t budget_bytes, int sort_by_norm,
                       int nthreads, uint64_t row_begin, uint64_t row_end, uint32_t* idx,
                       double* cos_out, double* norms, uint64_t* stats) {
  int rc = validate(x, n, d, k);
  if (rc) return rc;
  if (row_end > n || row_begin > row_end) return fail(ORACLE_E_INVALID, "bad row range");
  if (sort_by_norm && (row_begin != 0 || row_end != n)) return fail(ORACLE_E_INVALID, "row range needs sort_by_norm=0");
  set_threads(nthreads);
  int th_used = 1;
#ifdef _OPENMP
  th_used = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  oracle_row_norms64(x, n, d, norms, nthreads);

  /* knn.rs:413-418: stable ascending sort by norm (insertion of equal norms keeps index order) */
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * n);
  for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
  if (sort_by_norm) oracle_stable_sort_by_key(perm, n, norms);

  /* knn.rs:420-436: L2-normalized f32 working copy, zero rows stay zero */
  float* xs = (float*)calloc(n * (d ? d : 1), sizeof(float));
#pragma omp parallel for schedule(static)
  for (int64_t w = 0; w < (int64_t)n; ++w) {
    uint64_t o = perm[w];
    double nrm = norms[o];
    if (nrm > 0.0)
      for (uint64_t t = 0; t < d; ++t) xs[(uint64_t)w * d + t] = (float)((double)x[o * d + t] / nrm);
  }

  uint64_t rb, cb;
  oracle_block_sizes(n, row_block, col_block, budget_bytes, th_used, &rb, &cb);
  ctx_t c = {x, xs, norms, perm, n, d, k, (k + margin < n - 1 ? k + margin : n - 1), cb,
             ((double)d + 4.0) * (double)FLT_EPSILON_F64};
  uint64_t fallbacks = 0;
  uint64_t nblk = (row_end - row_begin + rb - 1) / rb;

#pragma omp parallel reduction(+ : fallbacks)
  {
    float* tile = (float*)malloc(sizeof(float) * rb * cb);
    heaps_t h;
    h.cap = c.cap;
    h.s = (float*)malloc(sizeof(float) * rb * c.cap);
    h.id = (uint32_t*)malloc(sizeof(uint32_t) * rb * c.cap);
    h.len = (uint64_t*)malloc(sizeof(uint64_t) * rb);
    uint64_t ccap = c.cap + 16;
    cand_t* cands = (cand_t*)malloc(sizeof(cand_t) * ccap);
#pragma omp for schedule(dynamic, 1)
    for (int64_t bi = 0; bi < (int64_t)nblk; ++bi) {
      uint64_t r0 = row_begin + (uint64_t)bi * rb;
      uint64_t rows = row_end - r0 < rb ? row_end - r0 : rb;
      memset(h.len, 0, sizeof(uint64_t) * rows);
      /* coarse_block / scan_block (knn.rs:603-610, 647-660): every column block, cosine never prunes */
      for (uint64_t c0 = 0; c0 < n; c0 += cb) {
        uint64_t cols = n - c0 < cb ? n - c0 : cb;
        oracle_tile_nt(xs + r0 * d, rows, xs + c0 * d, cols, d, tile);
        for (uint64_t r = 0; r < rows; ++r) scan_row(tile + r * cols, cols, c0, r0 + r, &h, r);
      }
      /* finalize_row (knn.rs:688-720) */
      for (uint64_t r = 0; r < rows; ++r) {
        uint64_t w_row = r0 + r, i = perm[w_row], nc = 0;
        const uint32_t* ids = h.id + r * c.cap;
        for (uint64_t t = 0; t < h.len[r]; ++t) push_rescored(&c, i, ids[t], cands, &nc);
        qsort(cands, nc, sizeof(cand_t), cand_cmp);
        if (c.cap < n - 1) {
          double floor_ = (double)heap_tau(&h, r);
          double s_k = cands[k - 1].score;
          if (s_k <= floor_ + c.eps) {
            ++fallbacks;
            gather_row(&c, w_row, s_k - c.eps, tile, &cands, &nc, &ccap);
            qsort(cands, nc, sizeof(cand_t), cand_cmp);
          }
        }
        for (uint64_t t = 0; t < k; ++t) {
          idx[w_row * k + t] = cands[t].j;
          cos_out[w_row * k + t] = cands[t].cos;
        }
      }
    }
    free(tile);
    free(h.s);
    free(h.id);
    free(h.len);
    free(cands);
  }
  free(xs);

  /* unpermute_rows (knn.rs:516-543); a plain out-of-place move restates the same mapping */
  if (sort_by_norm) {
    uint32_t* ti = (uint32_t*)malloc(sizeof(uint32_t) * n * k);
    double* tc = (double*)malloc(sizeof(double) * n * k);
    for (uint64_t w = 0; w < n; ++w) {
      memcpy(ti + (uint64_t)perm[w] * k, idx + w * k, sizeof(uint32_t) * k);
      memcpy(tc + (uint64_t)perm[w] * k, cos_out + w * k, sizeof(double) * k);
    }
    memcpy(idx, ti, sizeof(uint32_t) * n * k);
    memcpy(cos_out, tc, sizeof(double) * n * k);
    free(ti);
    free(tc);
  }
  free(perm);
  if (stats) {
    stats[0] = fallbacks;
    stats[1] = rb;
    stats[2] = cb;
    stats[3] = ((n + rb - 1) / rb) * ((n + cb - 1) / cb);
  }
  return 0;
}

/* ------------------------------------------------------------------ weights and union */

/* weight.rs:41-93.  weight: 0 tanimoto, 1 cosine, 2 jaccard. */
int oracle_reweight(uint64_t n, uint64_t k, const uint32_t* idx, const double* cos, const double* norms,
                    int weight, double* out) {
  for (uint64_t e = 0; e < n * k; ++e)
    if (idx[e] >= n) return fail(ORACLE_E_INVALID, "neighbor index %u out of range (n=%llu)", idx[e], (unsigned long long)n);
  if (weight == 0) {
    for (uint64_t e = 0; e < n * k; ++e) out[e] = oracle_tani_from_cos(norms[e / k], norms[idx[e]], cos[e]);
  } else if (weight == 1) {
    memcpy(out, cos, sizeof(double) * n * k);
  } else if (weight == 2) {
    uint32_t* sorted = (uint32_t*)malloc(sizeof(uint32_t) * n * k);
    memcpy(sorted, idx, sizeof(uint32_t) * n * k);
    for (uint64_t i = 0; i < n; ++i) oracle_sort_u32(sorted + i * k, k);
    for (uint64_t i = 0; i < n; ++i) {
      const uint32_t* a = sorted + i * k;
      for (uint64_t t = 0; t < k; ++t) {
        const uint32_t* b = sorted + (uint64_t)idx[i * k + t] * k;
        uint64_t p = 0, q = 0, s = 0;
        while (p < k && q < k) {
          if (a[p] < b[q]) ++p;
          else if (a[p] > b[q]) ++q;
          else { ++s; ++p; ++q; }
        }
        out[i * k + t] = (double)s / (2.0 * (double)k - (double)s);
      }
    }
    free(sorted);
  } else {
    return fail(ORACLE_E_INVALID, "unknown edge weight %d", weight);
  }
  return 0;
}

typedef struct {
  uint64_t key;
  double w;
} edge_t;

static int edge_cmp(const void* pa, const void* pb) {
  uint64_t a = ((const edge_t*)pa)->key, b = ((const edge_t*)pb)->key;
  return (a > b) - (a < b);
}

/* graph.rs:24-68: pair keys (lo<<32 | hi), sort, combine (a+b) - a*b, drop exact zeros.
 * Writes at most n*k edges; returns the edge count through *m. */
int oracle_fuzzy_union(uint64_t n, uint64_t k, const uint32_t* idx, const double* w, uint32_t* rows,
                       uint32_t* cols, double* wout, uint64_t* m) {
  edge_t* e = (edge_t*)malloc(sizeof(edge_t) * (n * k + 1));
  uint64_t ne = 0;
  for (uint64_t t = 0; t < n * k; ++t) {
    uint64_t i = t / k, j = idx[t];
    if (i == j) continue;
    uint64_t lo = i < j ? i : j, hi = i < j ? j : i;
    e[ne].key = (lo << 32) | hi;
    e[ne].w = w[t];
    ++ne;
  }
  qsort(e, ne, sizeof(edge_t), edge_cmp);
  uint64_t out = 0;
  for (uint64_t t = 0; t < ne;) {
    uint64_t key = e[t].key;
    double a = e[t].w, v;
    if (t + 1 < ne && e[t + 1].key == key) {
      double b = e[t + 1].w;
      v = (a + b) - a * b;
      t += 2;
    } else {
      v = a;
      t += 1;
    }
    if (v != 0.0) {
      rows[out] = (uint32_t)(key >> 32);
      cols[out] = (uint32_t)key;
      wout[out] = v;
      ++out;
    }
  }
  free(e);
  *m = out;
  return 0;
}

/* ------------------------------------------------------------------ CSR of the graph */

/* leiden.rs:52-100 (Csr::from_edges): an undirected edge list (each edge once) to compressed rows
 * holding both directions; exact-zero weights dropped; rows list neighbors in edge order.  The
 * first offending edge (in edge order, tested as the reference tests it: range, self-loop, weight)
 * fails the call.  indptr has n+1 entries; nbr and wout hold up to 2m entries. */
static void fmt_weight(char* buf, size_t len, double x) {
  if (x != x) { snprintf(buf, len, "NaN"); return; }
  if (x == 1.0 / 0.0) { snprintf(buf, len, "inf"); return; }
  if (x == -1.0 / 0.0) { snprintf(buf, len, "-inf"); return; }
  for (int p = 1; p <= 17; ++p) {
    snprintf(buf, len, "%.*g", p, x);
    if (strtod(buf, NULL) == x) return;
  }
}

int oracle_csr_from_edges(uint64_t n, uint64_t m, const uint32_t* rows, const uint32_t* cols, const double* w,
                          uint64_t* indptr, uint32_t* nbr, double* wout) {
  if (n > 0xffffffffull) return fail(ORACLE_E_INVALID, "%llu nodes exceed the u32 node index", (unsigned long long)n);
  for (uint64_t v = 0; v <= n; ++v) indptr[v] = 0;
  for (uint64_t e = 0; e < m; ++e) {
    const uint64_t i = rows[e], j = cols[e];
    const double x = w[e];
    if (i >= n || j >= n)
      return fail(ORACLE_E_INVALID, "edge %llu (%llu, %llu) is out of range (n=%llu)", (unsigned long long)e,
                  (unsigned long long)i, (unsigned long long)j, (unsigned long long)n);
    if (i == j) return fail(ORACLE_E_INVALID, "edge %llu is a self-loop on node %llu", (unsigned long long)e,
                            (unsigned long long)i);
    if (!(x >= 0.0 && x < 1.0 / 0.0)) {
      char b[64];
      fmt_weight(b, sizeof b, x);
      return fail(ORACLE_E_INVALID, "edge %llu has weight %s; weights must be finite and non-negative",
                  (unsigned long long)e, b);
    }
    if (x > 0.0) {
      ++indptr[i + 1];
      ++indptr[j + 1];
    }
  }
  for (uint64_t v = 0; v < n; ++v) indptr[v + 1] += indptr[v];
  uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * (n + 1));
  if (!fill) return fail(ORACLE_E_OOM, "out of memory");
  for (uint64_t v = 0; v < n; ++v) fill[v] = indptr[v];
  for (uint64_t e = 0; e < m; ++e) {
    if (!(w[e] > 0.0)) continue;
    const uint64_t i = rows[e], j = cols[e];
    nbr[fill[i]] = (uint32_t)j;
    wout[fill[i]++] = w[e];
    nbr[fill[j]] = (uint32_t)i;
    wout[fill[j]++] = w[e];
  }
  free(fill);
  return 0;
}
