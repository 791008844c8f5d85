// order.h -- locality order of the rows and the similarity bounds that let the fused scan skip
// candidate tiles (exact pruning for cosine selection).
//
// The reference scans every column block for cosine selection and prunes only for Tanimoto, by
// norm windows over a norm-sorted working order (crates/costal-core/src/knn.rs:413-436, 603-644,
// 663-678).  Here the working order groups similar rows instead: spherical k-means on a sample of
// the normalized rows, every row assigned to its nearest center, rows stably sorted by center.
// Every group of consecutive working rows (a query block, a candidate tile) gets a cone: the
// normalized mean of its rows and the largest angle from it to any of them.  For a query block Q
// and a candidate tile T with cone angles aQ, aT and center angle th, no pair (q, c) has a cosine
// above cos(max(0, th - aQ - aT)) (spherical triangle inequality).  The scan visits each block's
// tiles in decreasing bound order (up to bound buckets, each tile recording its bucket's largest
// bound) and stops at the first recorded bound below every query's admission
// threshold: the candidates left out score below every list floor, so the certificate of the
// rescoring (knn.rs:704-714) holds unchanged and the result is the brute force bit for bit.
//
// Tanimoto selection (T = c / (s - c), s = rho + 1/rho for the norm ratio rho) orders the rows of
// every direction group by norm as well, and bounds a pair of groups by c_max / (s_min - c_max),
// with c_max the cone bound and s_min the smallest s over the two groups' norm ranges (0 when
// c_max <= 0): the reference's norm window (knn.rs:246-282) and the cosine cones in one bound.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace costal {

// Rows of a candidate tile of the bound matrix (a whole number of the scan's LDS tiles): 256, or
// 512 where 256-row tiles would be more than the bound sort holds.
constexpr uint32_t PRUNE_TILE = 256;
// Largest tile count per query block the bound order handles (four tiles per thread of one
// 1024-thread workgroup).
constexpr uint32_t PRUNE_MAX_TILES = 4096;
// Rows of a rescan (rows that failed the first certificate) that get their own cones and bounds.
constexpr uint32_t RESCAN_PRUNE_ROWS = 8192;
// Bytes of LDS the k-means kernels hold per center feature (f32 center + i64 sum).
constexpr uint32_t KM_LDS_BYTES = 12;
constexpr uint32_t KM_LDS_MAX = 60 * 1024;
// Tanimoto: candidate tiles of rows per direction sub-center (whose rows are then sorted by norm),
// and the bits of the norm key (log-spaced over the rows' norm range)
constexpr uint32_t TANI_SUB_TILES = 4;
constexpr uint32_t TANI_NORM_BITS = 8;

struct OrderBufs {
  uint32_t* perm;    // n: working row -> original row
  uint32_t* iperm;   // n: original row -> working row
  float* centers;    // C x dp
  long long* sums;   // C x dp, fixed-point center sums (zero between calls)
  uint32_t* counts;  // C
  uint32_t* assign;  // n
  uint32_t* rank;    // n
  uint32_t* hist;    // (C + 1) x row blocks of 256, then the first working row of each
  uint32_t* totals;  // C + 1: rows per center, then the first working row of each
  uint32_t* sub_cnt; // C + 1: sub-centers of each first-level center (the zero rows: one)
  uint32_t* sub_off; // C + 1: the first sub-center of each
  float* sub_cen;    // G2 x dp
  uint32_t* key2;    // n: the sub-center of each working row, then sorted
  uint32_t* key2s;   // n
  uint32_t* perm2;   // n: radix-sort buffers
  uint32_t* rank2;   // n
  double* qcen;      // query blocks x dp: unit cone centers
  double* qang;      // query blocks x 3: cone angle, its cosine, its sine
  double* tcen;      // candidate tiles x dp
  double* tang;      // candidate tiles x 3
  double* qnr;       // Tanimoto: query blocks x 2: smallest and largest raw norm
  double* tnr;       // Tanimoto: candidate tiles x 2
  uint32_t* seq;     // query blocks x tiles: tile ids in decreasing (bucketed) bound order
  float* sb;         // query blocks x tiles: the bounds in that order
  // the same for the blocks of a rescan's row list (RESCAN_PRUNE_ROWS / qb blocks at most)
  double* rcen;
  double* rang;
  double* rnr;
  uint32_t* rseq;
  float* rsb;
};

struct OrderShape {
  uint32_t n, dp;      // rows, f32 working-copy width
  uint32_t C;          // k-means centers
  uint32_t sample;     // rows of the k-means sample (spread evenly over all rows)
  uint32_t iters;      // k-means iterations on the sample
  uint32_t G2;         // sub-centers at most, the zero rows' one included
  uint32_t sub_cap;    // sub-centers of one first-level center at most
  uint32_t sub_sample; // rows of a first-level center's sub-center fit at most
  uint32_t sub_iters;
  uint32_t key_bits;   // bits of a sort key: sub-center id, then (Tanimoto) the norm key
  uint32_t norm_bits;  // bits of the norm key (0: cosine selection)
  uint32_t tani;       // Tanimoto bounds
  uint32_t qb;         // rows of a query block
  uint32_t pt;         // rows of a candidate tile (PRUNE_TILE or twice that)
  uint32_t nblk, npt;  // query blocks, candidate tiles
};

// The default shape for n rows of dp features with query blocks of qb rows (tani: for Tanimoto
// selection; tile: rows of a candidate tile, 0 = PRUNE_TILE unless that makes more than
// PRUNE_MAX_TILES tiles).
OrderShape order_shape(uint32_t n, uint32_t dp, uint32_t qb, uint32_t clusters, bool tani = false,
                       uint32_t tile = 0);
// Whether the bound matrix fits the sort (and the scan has something to gain).
bool order_supported(const OrderShape& o);
// Device bytes of the buffers (each carved on 256-byte boundaries).
size_t order_bytes(const OrderShape& o);
// Carve the buffers out of `base` (order_bytes(o) bytes).
OrderBufs order_carve(void* base, const OrderShape& o);

// Working order of the rows of xs32 (normalized f32 rows, dp per row, original order): perm/iperm.
// Two levels: rows sorted by their nearest first-level center, then, within each first-level
// center's rows, by their nearest of its sub-centers (about one per candidate tile of rows,
// spherical k-means on a sample of those rows), so that consecutive rows, and so every query block
// and candidate tile, are tight groups.  Tanimoto: about one sub-center per TANI_SUB_TILES tiles,
// and within each, rows by norm (norms: the f64 row norms, in [nlo, nhi] where positive).
hipError_t order_rows(hipStream_t s, const float* xs32, const OrderShape& o, OrderBufs& b,
                      const double* norms = nullptr, double nlo = 0.0, double nhi = 0.0);
// Working copies in the working order, rows [n, n_pad) zero: xs32w (f32, dp32) and/or xs16w
// (FP16, dp16 per row, or hi then lo rows of dp16 when split16).  Either may be null.
hipError_t order_permute(hipStream_t s, const float* xs32, const OrderShape& o, const OrderBufs& b, uint32_t n_pad,
                         uint32_t dp16, int split16, float* xs32w, _Float16* xs16w);
// Cones (and for Tanimoto norm ranges, from norms) of the query blocks and candidate tiles, the
// bound matrix and its per-block order (by bound buckets; see bound_sort_kernel).
hipError_t order_bounds(hipStream_t s, const float* xs32, const OrderShape& o, OrderBufs& b,
                        const double* norms = nullptr);
// A rescan's row list (original rows, nr <= RESCAN_PRUNE_ROWS): sorted in place into working
// order (so that its blocks of qb rows are groups of similar rows), their cones, and every block's
// tiles in bound order in rseq / rsb.  Needs order_bounds first (the tile cones).
hipError_t order_rescan(hipStream_t s, const float* xs32, const OrderShape& o, OrderBufs& b, uint32_t* rows,
                        uint32_t nr, const double* norms = nullptr);

}  // namespace costal
