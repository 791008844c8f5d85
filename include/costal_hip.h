/*
 * costal_hip.h -- C ABI of the MI355X (gfx950) exact cosine kNN + Tanimoto edge weights.
 *
 * This is the drop-in boundary for CosTaL's graph-construction hot path.  Each entry point names
 * the reference interface it replaces (ike12th/CosTaL 2.1.0):
 *
 *   costal_knn_exact        crates/costal-core/src/knn.rs:360-379  knn_exact / knn_exact_with_stats
 *                           crates/costal-py/src/lib.rs:113-167     costal._costal._knn_exact
 *   costal_knn_bruteforce   crates/costal-core/src/knn.rs:879-924  knn_bruteforce[_select]
 *                           crates/costal-py/src/lib.rs:171-191     costal._costal._knn_bruteforce
 *   costal_reweight         crates/costal-core/src/weight.rs:41-60 weight::reweight
 *                           crates/costal-py/src/lib.rs:194-214     costal._costal._reweight
 *   costal_fuzzy_union      crates/costal-core/src/graph.rs:24-68  graph::fuzzy_union
 *                           crates/costal-py/src/lib.rs:218-239     costal._costal._fuzzy_union
 *   costal_csr_from_edges   crates/costal-core/src/leiden.rs:52-100 Csr::from_edges
 *
 * Conventions
 *   - Plain pointers and sizes only.  The *_device variants take device pointers and an optional
 *     hipStream_t (passed as void*, NULL = the library's own stream) and return when the work is
 *     complete; the plain variants take host pointers (row-major, C-contiguous) and copy.
 *   - The caller allocates every output: idx n*k (u32, original indices), cos n*k (f64),
 *     norms n (f64), weights n*k (f64).  Outputs are in original row order.
 *   - Every entry point returns 0 on success or one of the COSTAL_E_* codes below; the message is
 *     in costal_last_error() (thread-local), with the reference's Display text for KnnError
 *     (knn.rs:111-125).  No entry point aborts the process.
 *   - Results are identical for every tile shape, precision and thread layout: neighbor indices
 *     and cosines equal the f64 brute force bit for bit (knn.rs:884-924).
 *   - Thread-safe: calls on one device are serialized internally.
 */
#ifndef COSTAL_HIP_H
#define COSTAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* KnnError (knn.rs:102-109) plus the device-side failures */
#define COSTAL_OK 0
#define COSTAL_E_SHAPE 1          /* "data length {len} does not match shape {n}x{d}" */
#define COSTAL_E_BAD_K 2          /* "invalid k={k}: need 1 <= k <= n-1 (n={n})" */
#define COSTAL_E_TOO_MANY_ROWS 3  /* "n={n} exceeds the u32 index range" */
#define COSTAL_E_NON_FINITE 4     /* "row {row} contains NaN or inf" */
#define COSTAL_E_UNAVAILABLE 5    /* "matmul backend {name} is not available on this platform" */
#define COSTAL_E_DEVICE 6         /* HIP runtime error (no GPU, out of memory, launch failure) */
#define COSTAL_E_UNSUPPORTED 7    /* a size this build does not handle (e.g. k above 2048) */
#define COSTAL_E_INVALID 8        /* bad argument (null pointer, index out of range, bad enum) */

/* selection metric (knn.rs:62-86) */
#define COSTAL_SELECT_COSINE 0
#define COSTAL_SELECT_TANIMOTO 1

/* coarse-tile precision: plain FP32 MFMA; FP16 MFMA whose hits are re-scored in FP32 before
 * admission; or FP16 hi/lo split rows (3 FP16 MFMAs per slot) with an FP32-class error bound */
#define COSTAL_PREC_FP32 0
#define COSTAL_PREC_FP16 1
#define COSTAL_PREC_FP16X3 2

/* edge weights (weight.rs:10-26) */
#define COSTAL_WEIGHT_TANIMOTO 0
#define COSTAL_WEIGHT_COSINE 1
#define COSTAL_WEIGHT_JACCARD 2

/* KnnParams (knn.rs:26-60).  Zero-initialise, then set k; every other zero means "default". */
typedef struct costal_knn_params {
  uint64_t k;            /* neighbors per row, excluding the row itself */
  int64_t margin;        /* candidate list = min(k + margin, n - 1); negative = k (the default) */
  uint32_t select;       /* COSTAL_SELECT_* */
  uint32_t precision;    /* COSTAL_PREC_* */
  uint64_t row_block;    /* query rows per workgroup hint; 0 = tuned (speed only) */
  uint64_t col_block;    /* candidate columns per LDS tile hint; 0 = tuned (speed only) */
  uint32_t col_splits;   /* split the candidate range in this many parts; 0 = tuned (speed only) */
  /* KnnParams::prune, inverted so that zero is the default (on): 0 = the fused scan works on the
   * rows in a locality order (spherical k-means on a sample, rows sorted by nearest center; for
   * Tanimoto selection also by norm) and skips every candidate tile whose score bound (from the
   * cones, and norm ranges, of the query block and the tile) is below the admission threshold of
   * every list of the block; nonzero = every candidate tile is scanned in row order.  Speed only:
   * the result is the same bit for bit. */
  uint32_t no_prune;
  /* Optional per-row admission thresholds (f32, n entries; host memory for costal_knn_exact,
   * device memory for costal_knn_exact_device): candidates whose coarse similarity does not exceed
   * the row's threshold are never inserted.  Rows left with fewer than k admitted candidates, or
   * whose k-th rescored similarity is within the error bound of the threshold, are recomputed by
   * the plain scan, so the result stays exact whatever the thresholds are.  NULL = none. */
  const float* row_threshold;
  /* Optional learned thresholds (used when row_threshold is NULL): parameters of the threshold
   * model (costal_mlp_*; device memory for the device entry point, host memory otherwise),
   * evaluated on the GPU inside the call.  Row threshold = 1 - exp(model(row) + mlp_offset) minus
   * twice the candidate error bound.  The result stays exact whatever the model says. */
  const float* mlp_params;
  float mlp_offset;
  uint32_t mlp_bf16;     /* 1 = BF16 inference (products of bf16-rounded values, f32 sums) */
  uint32_t prune_clusters; /* k-means centers of the locality order; 0 = about n/3072 in [16, 256] */
  /* Second-chance offset of the learned thresholds (0 = none): rows left with fewer than k
   * candidates rescan with 1 - exp(model(row) + mlp_offset2) - margin when mlp_offset2 is above
   * mlp_offset (the rows short again go to the exact row scan); otherwise without a threshold. */
  float mlp_offset2;
  /* Rows of a candidate tile of the locality order (the unit the pruned scan skips): 256 or 512;
   * 0 = 256, or 512 when 256 would make more than 4096 tiles.  Speed only. */
  uint32_t prune_tile;
} costal_knn_params;

/* KnnStats (knn.rs:344-357) plus GPU-side counters */
typedef struct costal_knn_stats {
  uint64_t fallback_rows;     /* rows whose first scan failed the certificate and were recomputed */
  uint64_t row_block;         /* query rows per workgroup used */
  uint64_t col_block;         /* candidate columns per LDS tile used */
  uint64_t blocks_skipped;    /* (query block, candidate tile) pairs the pruned scan never read */
  uint64_t blocks_total;      /* query blocks x candidate tiles (x splits of a block's tiles: each
                                 pair is counted once) */
  uint64_t candidates;        /* K' = min(k + margin, n - 1) */
  uint64_t col_splits;        /* candidate-range splits used (2 for the balanced launch: a row's
                                 range in at most two pieces) */
  uint64_t short_rows;        /* rows with fewer than k admitted candidates (thresholds only) */
  uint64_t path;              /* 0 fused MFMA scan, 1 exact row scan (large K', tanimoto), 2 wide rows */
  uint64_t inserts;           /* candidate-list inserts made by the fused scan */
  uint64_t rowscan_rows;      /* rows that needed the exact full-row kernel */
  double device_ms;           /* device time of the call (hipEvent), 0 when not measured */
  double prep_ms;             /* norms + working copies kernel (hipEvent) */
  double coarse_ms;           /* fused MFMA tile + scan kernel (hipEvent) */
  double rescore_ms;          /* f64 rescoring + certificate kernel (hipEvent) */
  double fallback_ms;         /* full-row kernel (hipEvent) */
  /* Filled on COSTAL_E_NON_FINITE too: the row of KnnError::NonFinite{row} (knn.rs:102-127), the
   * first (lowest) row with a NaN or inf, and for a batch its sample; UINT64_MAX otherwise. */
  uint64_t nonfinite_row;
  uint64_t nonfinite_sample;
  /* Measured error of the candidate scores: the largest |list score - f64 cosine| over every
   * candidate admitted to a list in this call (first pass and rescans), next to the bound the
   * certificate used for them (knn.rs:164-188).  0 when no candidate list was built (row scan). */
  double coarse_err_max;
  double coarse_eps;
  double order_ms;            /* locality order, working copies and tile bounds of the pruned scan */
} costal_knn_stats;

/* Exact kNN by params->select from host memory.  norms may be NULL; weights (Tanimoto, n*k) may be
 * NULL.  Replaces knn_exact_with_stats (knn.rs:365-379). */
int costal_knn_exact(const float* x, uint64_t n, uint64_t d, const costal_knn_params* params,
                     uint32_t* idx, double* cos, double* norms, double* weights,
                     costal_knn_stats* stats);

/* The tile plan costal_knn_exact would use for this shape, without running it: fills path,
 * row_block, col_block, candidates, col_splits, blocks_total and coarse_eps of stats (the device's
 * CU count when a device is present, else 256; the first pass may still take the balanced launch,
 * which reports 2 splits).  Replaces block_sizes (knn.rs:551-569) as the place a caller reads the
 * blocking from; needs no GPU. */
int costal_knn_plan(uint64_t n, uint64_t d, const costal_knn_params* params, costal_knn_stats* stats);

/* The same from device memory; every pointer is a device pointer. */
int costal_knn_exact_device(const float* x, uint64_t n, uint64_t d,
                            const costal_knn_params* params, uint32_t* idx, double* cos,
                            double* norms, double* weights, costal_knn_stats* stats, void* stream);

/* ---- the same without allocating -------------------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  Arguments, results and
 * errors are those of costal_knn_exact_device, but every intermediate buffer (working copies,
 * candidate lists, counters, thresholds) lives in a static device array of COSTAL_SCRATCH_BYTES
 * per GPU, held from the library's set-up until the process ends; nothing is allocated or freed.
 * A call whose buffers do not fit, even with one candidate range per row, fails with
 * COSTAL_E_UNSUPPORTED and a message giving the bytes it needs.  Calls on one GPU share that
 * GPU's scratch: the library serializes them and each returns only when its work is complete, so
 * two calls never use it at once (two copies of the library in one process would race). */
#define COSTAL_SCRATCH_BYTES (64u << 20)
int costal_knn_exact_device_noalloc(const float* x, uint64_t n, uint64_t d,
                                    const costal_knn_params* params, uint32_t* idx, double* cos,
                                    double* norms, double* weights, costal_knn_stats* stats,
                                    void* stream);

/* ---- batched samples -----------------------------------------------------------------------
 * No single counterpart in the reference (it runs one knn() per sample, knn.rs:360-379): many
 * independent samples of uneven size in one pass, each with exactly the result costal_knn_exact
 * gives for it alone.  One persistent scan launch pulls (sample, query block, candidate split)
 * work items from a queue; the whole call is captured once as a hipGraph and replayed while the
 * batch layout (pointers, sizes, parameters) stays the same.  Cosine selection, d <= 64. */
typedef struct costal_sample {
  const float* x;        /* device, n x d row-major */
  uint64_t n, d;
  uint32_t* idx;         /* device outputs: n*k, n*k, n, n*k (weights may be NULL) */
  double* cos;
  double* norms;
  double* weights;
  int64_t margin;        /* candidate list min(k + margin, n - 1); negative = params->margin */
  uint32_t col_splits;   /* candidate-range splits of this sample; 0 = planned */
  uint32_t reserved;
} costal_sample;

/* params->k, ->precision and ->margin apply to every sample (select must be cosine). */
int costal_knn_batched_device(const costal_sample* samples, uint32_t nsamples,
                              const costal_knn_params* params, costal_knn_stats* stats, void* stream);

/* Brute force by the definition: every j != i scored in f64, ranked by (score desc, index asc).
 * Replaces knn_bruteforce_select (knn.rs:884-924). */
int costal_knn_bruteforce(const float* x, uint64_t n, uint64_t d, uint64_t k, uint32_t select,
                          uint32_t* idx, double* cos, double* norms);

/* Edge weights on an n x k kNN (weight.rs:41-93); every idx must be < n (lib.rs:66-78). */
int costal_reweight(uint64_t n, uint64_t k, const uint32_t* idx, const double* cos,
                    const double* norms, uint32_t weight, double* out);
int costal_reweight_device(uint64_t n, uint64_t k, const uint32_t* idx, const double* cos,
                           const double* norms, uint32_t weight, double* out, void* stream);

/* Fuzzy union S + S^T - S.S^T of an n x k weighted kNN (graph.rs:24-68): the upper-triangular
 * edge list in ascending (row, col) order with self-loops and exact zeros dropped.  rows, cols and
 * w must hold n*k entries; the edge count is written to *m. */
int costal_fuzzy_union(uint64_t n, uint64_t k, const uint32_t* idx, const double* w,
                       uint32_t* rows, uint32_t* cols, double* wout, uint64_t* m);

/* The compressed rows the CPU Leiden consumes (leiden.rs:52-100, Csr::from_edges) from an edge list
 * holding each undirected edge once (costal_fuzzy_union's output): both directions of every edge
 * with a positive weight, rows listing neighbors in edge order (ascending for a sorted
 * upper-triangular list); exact-zero weights dropped.  indptr holds n+1 entries, nbr and wout up
 * to 2m; the entry count is written to *nnz.  Edges out of range, self-loops and negative or
 * non-finite weights fail with the reference's message for the first such edge. */
int costal_csr_from_edges(uint64_t n, uint64_t m, const uint32_t* rows, const uint32_t* cols,
                          const double* w, uint64_t* indptr, uint32_t* nbr, double* wout,
                          uint64_t* nnz);

/* ---- learned admission threshold ----------------------------------------------------------
 * No counterpart in the reference: the per-row threshold model of the fused scan.  A 3-layer MLP
 * d -> 64 -> 64 -> 1 (ReLU) on L2-normalized rows predicting u = log(1 - s_k + 1e-6) of the row's
 * k-th neighbor cosine s_k.  Parameters are one flat f32 array:
 *   W1[64][d], b1[64], W2[64][64], b2[64], w3[64], b3   (costal_mlp_param_count(d) floats).
 * All pointers are device pointers; stream as above.  Training is deterministic. */
#define COSTAL_MLP_HIDDEN 64
#define COSTAL_MLP_MAX_FEATURES 64
uint64_t costal_mlp_param_count(uint64_t d);
/* floats of activation scratch a gradient call needs for n rows */
uint64_t costal_mlp_act_floats(uint64_t n);
/* out[r] = 1 - exp(model(x_r) + offset) - margin, or model(x_r) - offset when raw != 0 */
int costal_mlp_predict_device(const float* x, uint64_t n, uint64_t d, const float* params, uint32_t bf16,
                              float offset, float margin, uint32_t raw, float* out, void* stream);
/* gradient of the mean squared error of model(x) against target over the n rows of a batch;
 * act holds costal_mlp_act_floats(n) floats; loss (1 float) may be NULL */
int costal_mlp_grad_device(const float* x, const float* target, uint64_t n, uint64_t d,
                           const float* params, float* act, float* grad, float* loss, void* stream);
/* one Adam step (step counts from 1) */
int costal_mlp_adam_device(uint64_t count, float* params, const float* grad, float* m, float* v,
                           float lr, float beta1, float beta2, float eps, uint64_t step, void* stream);
const char* costal_mlp_last_error(void);

/* Self-check of the matrix core (no counterpart in the reference): eight MFMA results into out
 * (host memory, 8 floats) that show whether FP32 accumulation rounds to nearest and whether FP16
 * subnormal inputs are kept, the two properties the split-FP16 candidate bound relies on.
 * Expected on round-to-nearest hardware that keeps subnormals:
 *   out[0] = 1 + 2^-23, out[1] = 1 + 2^-23 (sum rounded once) or 1 (rounded per addition),
 *   out[2] = 1 - 2^-24, out[3] = out[4] = out[7] = 2^-24, out[5] = 1 + 2^-23, out[6] = 1 - 2^-24. */
int costal_mfma_probe(float* out);

/* Thread-local message of the last failing call on this thread. */
const char* costal_last_error(void);
/* Library version string. */
const char* costal_version(void);
/* Human-readable device description ("gfx950, 256 CUs, ...") into buf; 0 on success. */
int costal_device_info(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
