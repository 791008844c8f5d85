/*
 * costal_oracle.h -- CPU restatement of CosTaL's exact kNN path.  TEST INFRASTRUCTURE ONLY
 * (tests/, __graft_entry__.smoke(), bench.py's CPU-baseline leg); never part of the product path.
 */
#ifndef COSTAL_ORACLE_H
#define COSTAL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* KnnError codes (crates/costal-core/src/knn.rs:102-127), same numbering as include/costal_hip.h */
#define ORACLE_E_SHAPE 1
#define ORACLE_E_BAD_K 2
#define ORACLE_E_TOO_MANY_ROWS 3
#define ORACLE_E_NON_FINITE 4
#define ORACLE_E_INVALID 8
#define ORACLE_E_OOM 9

/* f32::EPSILON = 2^-23, as an f64 */
#define FLT_EPSILON_F64 1.1920928955078125e-07

const char* oracle_last_error(void);
double oracle_dot64(const float* a, const float* b, uint64_t d);
double oracle_tani_from_cos(double a, double b, double cos);
int oracle_row_norms64(const float* x, uint64_t n, uint64_t d, double* norms, int nthreads);
int oracle_knn_bruteforce(const float* x, uint64_t n, uint64_t d, uint64_t k, int select_tanimoto,
                          uint32_t* idx, double* cos, double* norms, int nthreads);
void oracle_block_sizes(uint64_t n, uint64_t row_block, uint64_t col_block, uint64_t budget_bytes,
                        int nthreads, uint64_t* rb, uint64_t* cb);
int oracle_knn_blocked(const float* x, uint64_t n, uint64_t d, uint64_t k, uint64_t margin,
                       uint64_t row_block, uint64_t col_block, uint64_t budget_bytes, int sort_by_norm,
                       int nthreads, uint64_t row_begin, uint64_t row_end, uint32_t* idx,
                       double* cos, double* norms, uint64_t* stats);
uint64_t oracle_heap_keep(const float* scores, const uint32_t* ids, uint64_t m, uint64_t cap,
                          uint32_t* kept);
int oracle_reweight(uint64_t n, uint64_t k, const uint32_t* idx, const double* cos,
                    const double* norms, int weight, double* out);
int oracle_fuzzy_union(uint64_t n, uint64_t k, const uint32_t* idx, const double* w,
                       uint32_t* rows, uint32_t* cols, double* wout, uint64_t* m);

int oracle_csr_from_edges(uint64_t n, uint64_t m, const uint32_t* rows, const uint32_t* cols,
                          const double* w, uint64_t* indptr, uint32_t* nbr, double* wout);

/* oracle_gemm.c */
void oracle_tile_nt(const float* a, uint64_t m, const float* b, uint64_t nb, uint64_t d, float* out);
void oracle_stable_sort_by_key(uint32_t* perm, uint64_t n, const double* key);
void oracle_sort_u32(uint32_t* v, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
