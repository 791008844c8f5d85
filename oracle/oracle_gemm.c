/*
 * oracle_gemm.c -- the f32 tile product and small sorting helpers of the CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see costal_oracle.c).
 *
 * oracle_tile_nt restates the semantics of matmul::tile_nt (crates/costal-core/src/matmul.rs:52-106):
 * out[m x nb] = a[m x d] . b[nb x d]^T, row stride nb.  Its rounding is deliberately unpinned (the
 * reference's own backends differ too); only the coarse stage uses it and the certificate bounds its
 * error by (d+4)*2^-23.  This file may be compiled with FP contraction and vectorization.
 */
#include <stdlib.h>
#include <string.h>

#include "costal_oracle.h"

#define NB_CHUNK 256

void oracle_tile_nt(const float* a, uint64_t m, const float* b, uint64_t nb, uint64_t d, float* out) {
  if (m == 0 || nb == 0) return;
  if (d == 0) {
    memset(out, 0, sizeof(float) * m * nb);
    return;
  }
  /* transpose a chunk of b to [d][chunk] so the inner loop is a unit-stride axpy over columns */
  float* bt = (float*)malloc(sizeof(float) * d * NB_CHUNK);
  for (uint64_t j0 = 0; j0 < nb; j0 += NB_CHUNK) {
    uint64_t nc = nb - j0 < NB_CHUNK ? nb - j0 : NB_CHUNK;
    for (uint64_t j = 0; j < nc; ++j)
      for (uint64_t t = 0; t < d; ++t) bt[t * NB_CHUNK + j] = b[(j0 + j) * d + t];
    uint64_t i = 0;
    for (; i + 4 <= m; i += 4) {
      float acc0[NB_CHUNK], acc1[NB_CHUNK], acc2[NB_CHUNK], acc3[NB_CHUNK];
      for (uint64_t j = 0; j < nc; ++j) acc0[j] = acc1[j] = acc2[j] = acc3[j] = 0.0f;
      const float *a0 = a + i * d, *a1 = a0 + d, *a2 = a1 + d, *a3 = a2 + d;
      for (uint64_t t = 0; t < d; ++t) {
        const float* br = bt + t * NB_CHUNK;
        float v0 = a0[t], v1 = a1[t], v2 = a2[t], v3 = a3[t];
        for (uint64_t j = 0; j < nc; ++j) {
          float bv = br[j];
          acc0[j] += v0 * bv;
          acc1[j] += v1 * bv;
          acc2[j] += v2 * bv;
          acc3[j] += v3 * bv;
        }
      }
      memcpy(out + (i + 0) * nb + j0, acc0, sizeof(float) * nc);
      memcpy(out + (i + 1) * nb + j0, acc1, sizeof(float) * nc);
      memcpy(out + (i + 2) * nb + j0, acc2, sizeof(float) * nc);
      memcpy(out + (i + 3) * nb + j0, acc3, sizeof(float) * nc);
    }
    for (; i < m; ++i) {
      float acc[NB_CHUNK];
      for (uint64_t j = 0; j < nc; ++j) acc[j] = 0.0f;
      const float* ar = a + i * d;
      for (uint64_t t = 0; t < d; ++t) {
        const float* br = bt + t * NB_CHUNK;
        float v = ar[t];
        for (uint64_t j = 0; j < nc; ++j) acc[j] += v * br[j];
      }
      memcpy(out + i * nb + j0, acc, sizeof(float) * nc);
    }
  }
  free(bt);
}

/* Stable ascending sort of perm by key[perm[.]] (knn.rs:415-418; key order = f64 total order,
 * and norms are non-negative and finite, so '<' agrees with it).  Merge sort: stable by construction. */
void oracle_stable_sort_by_key(uint32_t* perm, uint64_t n, const double* key) {
  uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint64_t width = 1; width < n; width *= 2) {
    for (uint64_t lo = 0; lo < n; lo += 2 * width) {
      uint64_t mid = lo + width < n ? lo + width : n;
      uint64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
      uint64_t p = lo, q = mid, o = lo;
      while (p < mid && q < hi) tmp[o++] = (key[perm[q]] < key[perm[p]]) ? perm[q++] : perm[p++];
      while (p < mid) tmp[o++] = perm[p++];
      while (q < hi) tmp[o++] = perm[q++];
    }
    memcpy(perm, tmp, sizeof(uint32_t) * n);
  }
  free(tmp);
}

static int u32_cmp(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return (x > y) - (x < y);
}

void oracle_sort_u32(uint32_t* v, uint64_t n) { qsort(v, n, sizeof(uint32_t), u32_cmp); }
