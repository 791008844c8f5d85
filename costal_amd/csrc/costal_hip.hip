// costal_hip.hip -- exact cosine kNN + Tanimoto edge weights for MI355X (gfx950).
//
// The path (reference = ike12th/CosTaL 2.1.0, crates/costal-core/src/knn.rs):
//
//   prep_kernel      row norms in f64 with a sequential sum (knn.rs:289-296, 315-323), the
//                    L2-normalized f32 working copy (knn.rs:420-436) and its FP16 copy, and the
//                    first non-finite row (knn.rs:325-342).
//   order.hip        the locality order of the rows and each query block's candidate tiles in
//                    decreasing bound order (the pruned scan; Tanimoto: norm ranges as well).
//   coarse_kernel    (scan.h) the fused blocked coarse selection (knn.rs:603-660, 759-875): MFMA
//                    similarity tiles of 32 candidates x 32 queries, staged through LDS, never
//                    written to HBM; each lane keeps a sorted list of its best candidates in
//                    registers (two lists per query), FP32, FP16 (re-scored in f32 on admission)
//                    or split-FP16 tiles; cosine or Tanimoto selection.
//   rescore_kernel   finalize_row (knn.rs:688-720): f64 rescoring over the original rows,
//                    ranking by (score desc, index asc), the epsilon certificate; rows that fail it
//                    are rescanned, then queued for the full-row path.  Fuses the Tanimoto weight
//                    (weight.rs:28-37).
//   rowscan_kernel   gather_row (knn.rs:723-743) and the brute force (knn.rs:884-924): one workgroup
//                    streams every candidate of a row, keeps an exact running top-k in LDS.
//   reweight_kernel  weight::reweight (weight.rs:41-60).
//
// Exactness: every admitted candidate is re-scored in f64 exactly as the reference does; a
// candidate that was never admitted has coarse score <= the row's floor, hence exact score
// <= floor + eps; a row whose k-th f64 score is not strictly above floor + eps is recomputed by the
// full-row path.  So idx/cos equal the f64 brute force bit for bit for any tile shape, precision or
// admission threshold.
//
// Built with -ffp-contract=off: the f64 expressions (cos64, tani_from_cos) must round as written.
// f32 inner products that only feed the coarse stage use explicit fmaf.  Device code is built
// without IEEE NaN quieting (inputs are checked for non-finite values before any scan runs), so
// max/compare chains need no canonicalizing moves, and MFMA results stay in VGPRs.

#include <hip/hip_runtime.h>

#include "order.h"
#include "scan.h"

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/costal_hip.h"

#define COSTAL_VERSION "costal-mi355x 0.2.0 (gfx950)"

// The workspace of costal_knn_exact_device_noalloc: working copies, candidate lists, counters and
// thresholds of one call, COSTAL_SCRATCH_BYTES per device, held by every GPU the library runs on
// until the process ends.  Host code never uses this name as an address: device_init looks the
// device's copy up with hipGetSymbolAddress and keeps that pointer per device.
static __device__ __align__(256) uint8_t scratch[COSTAL_SCRATCH_BYTES];


namespace {


// ------------------------------------------------------------------------------------------------
// Scalar f64 primitives (bit-exact with the reference)

// knn.rs:289-296: f64 sum in index order.  f32*f32 is exact in f64, so mul+add == fma here.
__device__ __forceinline__ double dot64(const float* __restrict__ a, const float* __restrict__ b,
                                        uint32_t d) {
  double s = 0.0;
#pragma unroll 8
  for (uint32_t t = 0; t < d; ++t) s += (double)a[t] * (double)b[t];
  return s;
}

// dot64 over rows of rs floats.  V4: rs is a multiple of 4, the rows are 16-byte aligned and hold
// zeros past their features; they are read 16 bytes at a time (a gathered row then costs one
// vector load per 4 features instead of one per feature, and never straddles a cache line more
// than the row's own length needs).  The sum starts at +0.0 and so is never -0.0 (x + -x rounds
// to +0.0), and the padding adds +0.0 products, which leave it unchanged: bit-identical to dot64.
// Two 16-byte words of each row in flight: the rescoring kernels then keep to 64 VGPRs (eight
// waves per SIMD; four words took 95 and made 10^6 x 40 rescoring 5% slower, one word 34%).
template <bool V4>
__device__ __forceinline__ double dot64r(const float* __restrict__ a, const float* __restrict__ b, uint32_t rs) {
  if constexpr (V4) {
    const f32x4* __restrict__ a4 = reinterpret_cast<const f32x4*>(a);
    const f32x4* __restrict__ b4 = reinterpret_cast<const f32x4*>(b);
    double s = 0.0;
#pragma unroll 2
    for (uint32_t c = 0; c < rs / 4; ++c) {
      const f32x4 u = a4[c], v = b4[c];
      s += (double)u[0] * (double)v[0];
      s += (double)u[1] * (double)v[1];
      s += (double)u[2] * (double)v[2];
      s += (double)u[3] * (double)v[3];
    }
    return s;
  } else {
    return dot64(a, b, rs);
  }
}

// knn.rs:300-312
__device__ __forceinline__ double cos64(double dot, double ni, double nj) {
  if (ni == 0.0 || nj == 0.0) return 0.0;
  double c = dot / (ni * nj);
  return c == 0.0 ? 0.0 : c;
}

// weight.rs:28-37 (evaluation order kept: (a*b)*cos, then ((a*a)+(b*b))-d)
__device__ __forceinline__ double tani_from_cos(double a, double b, double c) {
  double d = a * b * c;
  if (d == 0.0) return 0.0;
  return d / (a * a + b * b - d);
}

// The selection score of a pair from its f64 dot product: the cosine, or the Tanimoto score of
// that cosine (Selector::rescore, knn.rs:152-153, 257-260)
__device__ __forceinline__ double score64(double dot, double ni, double nj, bool tani) {
  const double c = cos64(dot, ni, nj);
  return tani ? tani_from_cos(ni, nj, c) : c;
}


// (score desc, index asc): the order of sort_cands (knn.rs:589-591).  Scores are never NaN and
// -0.0 never occurs (cos64 normalizes it), so '>' agrees with total_cmp.
__device__ __forceinline__ bool ranks_before(double sa, uint32_t ia, double sb, uint32_t ib) {
  return sa > sb || (sa == sb && ia < ib);
}

// ------------------------------------------------------------------------------------------------
// prep: norms, normalized copies, non-finite check.  One thread per row; rows in [n, n_pad) are
// written as zeros so that full tiles can be read without bounds checks.

// prep with coalesced loads, 32 rows per workgroup.  The rows come through LDS in coalesced
// 64-feature chunks (the next chunk in flight while the current one is summed), and thread r sums
// row r's squares in feature order (f64, index order: bit-identical norms), flagging the first
// row with a NaN or inf (the test reads the bits as integers); the normalized copies are written
// straight from global, coalesced: the f32 copy f32(f64(x)/norm), the FP16 copy, or the split
// one (hi = fp16(v), lo = fp16(v - hi), v - hi exact in f32; hi then lo halves per row), and when
// xr is given the rows themselves padded with zeros to dp32 floats (what the rescoring reads).  Rows in
// [n, n_pad) are zeros, so that whole tiles can be read without bounds checks.  (One thread per
// row reading its own row took 1.3 ms for 10^6 rows of 40 features.)
constexpr int PREP_ROWS = 32;
constexpr int PREP_CF = 64;

__global__ __launch_bounds__(256) void prep_kernel(const float* __restrict__ x, uint32_t n, uint32_t d,
                                                        uint32_t n_pad, uint32_t dp32, uint32_t dp16, int split16,
                                                        double* __restrict__ norms, float* __restrict__ xs32,
                                                        _Float16* __restrict__ xs16, float* __restrict__ xr,
                                                        uint32_t* __restrict__ bad_row) {
  constexpr int PER = PREP_ROWS * PREP_CF / 256;  // elements per thread per chunk
  __shared__ uint32_t tile[2][PREP_ROWS][PREP_CF + 1];  // the rows' bits (the non-finite test reads bits)
  __shared__ double nrm_s[PREP_ROWS];
  const uint32_t t = tid_x(), r0 = bid_x() * PREP_ROWS;
  const uint32_t nr = r0 < n ? (n - r0 < (uint32_t)PREP_ROWS ? n - r0 : (uint32_t)PREP_ROWS) : 0;
  const uint32_t* xb = reinterpret_cast<const uint32_t*>(x);
  const uint32_t nch = (d + PREP_CF - 1) / PREP_CF;
  uint32_t reg[PER];
  auto load = [&](uint32_t c0) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const uint32_t e = t + 256 * m, r = e / PREP_CF, f = e % PREP_CF;
      reg[m] = (r < nr && c0 + f < d) ? xb[(size_t)(r0 + r) * d + c0 + f] : 0u;
    }
  };
  auto store = [&](int b) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const uint32_t e = t + 256 * m;
      tile[b][e / PREP_CF][e % PREP_CF] = reg[m];
    }
  };
  double s = 0.0;
  bool bad = false;
  load(0);
  store(0);
  __syncthreads();
  for (uint32_t c = 0; c < nch; ++c) {
    if (c + 1 < nch) load((c + 1) * PREP_CF);
    if (t < nr) {
      const uint32_t c0 = c * PREP_CF;
      const uint32_t m = d - c0 < (uint32_t)PREP_CF ? d - c0 : (uint32_t)PREP_CF;
      for (uint32_t f = 0; f < m; ++f) {
        const uint32_t b = tile[c & 1][t][f];
        bad |= (b & 0x7f800000u) == 0x7f800000u;  // inf or NaN
        const float v = __uint_as_float(b);
        s += (double)v * (double)v;
      }
    }
    if (c + 1 < nch) store((c + 1) & 1);
    __syncthreads();
  }
  if (t < PREP_ROWS) {
    const double nrm = sqrt(s);
    nrm_s[t] = t < nr ? nrm : 0.0;
    if (t < nr) {
      norms[r0 + t] = nrm;
      if (bad) atomicMin(bad_row, r0 + t);
    }
  }
  __syncthreads();
  for (uint32_t e = t; e < (uint32_t)PREP_ROWS * dp32; e += 256) {
    const uint32_t r = e / dp32, f = e % dp32;
    if (r0 + r >= n_pad) break;
    const double nr_ = nrm_s[r];
    const bool in = r < nr && f < d;
    const float v = in ? __uint_as_float(xb[(size_t)(r0 + r) * d + f]) : 0.0f;
    xs32[(size_t)(r0 + r) * dp32 + f] = (in && nr_ > 0.0) ? (float)((double)v / nr_) : 0.0f;
    if (xr) xr[(size_t)(r0 + r) * dp32 + f] = v;
  }
  if (xs16) {
    // the FP16 copy (split: hi then lo halves), one feature per thread
    const uint32_t s16 = split16 ? 2 * dp16 : dp16;
    for (uint32_t e = t; e < (uint32_t)PREP_ROWS * dp16; e += 256) {
      const uint32_t r = e / dp16, f = e % dp16;
      if (r0 + r >= n_pad) break;
      const double nr_ = nrm_s[r];
      const float v =
          (r < nr && f < d && nr_ > 0.0) ? (float)((double)__uint_as_float(xb[(size_t)(r0 + r) * d + f]) / nr_) : 0.0f;
      const _Float16 hi = (_Float16)v;
      xs16[(size_t)(r0 + r) * s16 + f] = hi;
      if (split16) xs16[(size_t)(r0 + r) * s16 + dp16 + f] = (_Float16)(v - (float)hi);
    }
  }
}


// ------------------------------------------------------------------------------------------------
// Wide rows (more features than the fused scan keeps in registers): two kernels per block of query
// rows.  wide_gemm_kernel computes the f32 scores of the block's queries against every candidate
// with FP32 MFMAs over 32-feature chunks staged through LDS, and stores each 32 x 32 block of
// scores in the accumulator layout (lane-major, 4 KB per block); wide_scan_kernel reads the blocks
// back, one 64-byte run per lane, and keeps the same per-lane candidate lists as the fused scan, so
// the rescoring, the certificate and the fallbacks are shared.  The f32 working copy has rows of a
// multiple of 32 features (zero padded) for this path.
//
// Score block (qs, cs) = queries qs*32 .. +32 of the block, candidates cs*32 .. +32, at floats
// ((qs * ncs) + cs) * 1024; lane l holds query l % 32 and candidates 8*(v/4) + 4*(l/32) + (v%4) at
// l*16 + v, v = 0..15.

constexpr int WIDE_KC = 32;   // features per chunk: J = 4 sixteen-byte slots per lane half
constexpr int WIDE_QT = 64;   // queries per gemm workgroup (2 sub-tiles)
constexpr int WIDE_CT = 128;  // candidates per gemm workgroup (4 sub-tiles)
constexpr int WIDE_QB = 128;  // queries per scan workgroup (4 waves)

__global__ __launch_bounds__(256) void wide_gemm_kernel(const float* __restrict__ xs32, uint32_t dp, uint32_t q0,
                                                        uint32_t ncs, float* __restrict__ S) {
  constexpr int J = WIDE_KC / 8;
  constexpr int ROWB = WIDE_KC * 4;
  __shared__ __attribute__((aligned(16))) uint8_t lq[2][WIDE_QT * ROWB];
  __shared__ __attribute__((aligned(16))) uint8_t lc[2][WIDE_CT * ROWB];
  const int tid = tid_x(), w = tid >> 6, l = tid & 63, h = l >> 5, l32 = l & 31;
  const int wq = w & 1, wc = w >> 1;  // the wave's query sub-tile and candidate pair of sub-tiles
  const uint8_t* qsrc = reinterpret_cast<const uint8_t*>(xs32 + (size_t)(q0 + bid_x() * WIDE_QT) * dp);
  const uint8_t* csrc = reinterpret_cast<const uint8_t*>(xs32 + (size_t)(bid_y() * WIDE_CT) * dp);
  const uint32_t rowbytes = dp * 4, nchunk = dp / WIDE_KC;
  // LDS slot e = ((sub*J + j)*2 + half)*32 + r holds 16-byte slot 2j + half of the chunk of row
  // sub*32 + r; thread t stages slots t + 256 i (conflict-free stores and lane reads)
  auto src_off = [&](int e, uint32_t c) -> uint32_t {
    const int r = e & 31, hh = (e >> 5) & 1, j = (e >> 6) % J, sub = (e >> 6) / J;
    return (uint32_t)(sub * 32 + r) * rowbytes + c * ROWB + (uint32_t)(2 * j + hh) * 16;
  };
  uint4 rq0, rq1, rc0, rc1, rc2, rc3;
  auto load = [&](uint32_t c) __attribute__((always_inline)) {
    rq0 = *reinterpret_cast<const uint4*>(qsrc + src_off(tid, c));
    rq1 = *reinterpret_cast<const uint4*>(qsrc + src_off(tid + 256, c));
    rc0 = *reinterpret_cast<const uint4*>(csrc + src_off(tid, c));
    rc1 = *reinterpret_cast<const uint4*>(csrc + src_off(tid + 256, c));
    rc2 = *reinterpret_cast<const uint4*>(csrc + src_off(tid + 512, c));
    rc3 = *reinterpret_cast<const uint4*>(csrc + src_off(tid + 768, c));
  };
  auto store = [&](int b) __attribute__((always_inline)) {
    reinterpret_cast<uint4*>(lq[b])[tid] = rq0;
    reinterpret_cast<uint4*>(lq[b])[tid + 256] = rq1;
    reinterpret_cast<uint4*>(lc[b])[tid] = rc0;
    reinterpret_cast<uint4*>(lc[b])[tid + 256] = rc1;
    reinterpret_cast<uint4*>(lc[b])[tid + 512] = rc2;
    reinterpret_cast<uint4*>(lc[b])[tid + 768] = rc3;
  };
  f32x16 acc0 = {}, acc1 = {};
  auto compute = [&](int b) __attribute__((always_inline)) {
    const uint4* Q = reinterpret_cast<const uint4*>(lq[b]);
    const uint4* Cc = reinterpret_cast<const uint4*>(lc[b]);
    uint4 q[J], b0[J], b1[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      q[j] = Q[((wq * J + j) * 2 + h) * 32 + l32];
      b0[j] = Cc[(((2 * wc) * J + j) * 2 + h) * 32 + l32];
      b1[j] = Cc[(((2 * wc + 1) * J + j) * 2 + h) * 32 + l32];
    }
    acc0 = tile_mfma<PREC_F32, J>(b0, q, acc0);
    acc1 = tile_mfma<PREC_F32, J>(b1, q, acc1);
  };
  load(0);
  store(0);
  __syncthreads();
  for (uint32_t c = 0; c < nchunk; ++c) {
    if (c + 1 < nchunk) load(c + 1);
    compute(c & 1);
    if (c + 1 < nchunk) store((c + 1) & 1);
    __syncthreads();
  }
  const uint32_t qs = bid_x() * 2 + wq, cs = bid_y() * 4 + 2 * wc;
  float4* o0 = reinterpret_cast<float4*>(S + ((size_t)qs * ncs + cs) * 1024 + l * 16);
  float4* o1 = reinterpret_cast<float4*>(S + ((size_t)qs * ncs + cs + 1) * 1024 + l * 16);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    o0[v] = make_float4(acc0[4 * v], acc0[4 * v + 1], acc0[4 * v + 2], acc0[4 * v + 3]);
    o1[v] = make_float4(acc1[4 * v], acc1[4 * v + 1], acc1[4 * v + 2], acc1[4 * v + 3]);
  }
}

// The per-lane lists of the fused scan over stored score blocks: queries q0 + [0, nq) (the block),
// candidate sub-tiles of split bid_y(); lists go to the slots of the queries' rows.
template <int KPR>
__global__ __launch_bounds__(256) void wide_scan_kernel(
    const float* __restrict__ S, uint32_t q0, uint32_t nq, uint32_t ncs, uint32_t n, int kp, uint32_t nsplit,
    uint32_t cs_per_split, uint32_t* __restrict__ cand_idx, float* __restrict__ cand_sc,
    uint32_t* __restrict__ cand_cnt, float* __restrict__ cand_floor, unsigned long long* __restrict__ inserts) {
  const int tid = tid_x(), w = tid >> 6, l = tid & 63, h = l >> 5, l32 = l & 31;
  const uint32_t qs = bid_x() * 4 + w;
  const uint32_t ql = qs * 32 + l32;
  const bool qvalid = ql < nq;
  const uint32_t qg = q0 + (qvalid ? ql : 0);
  const uint32_t split = bid_y();
  const uint32_t cs0 = split * cs_per_split;
  const uint32_t cs1 = cs0 + cs_per_split < ncs ? cs0 + cs_per_split : ncs;
  const uint32_t c_end = cs1 * 32 < n ? cs1 * 32 : n;
  float ls[KPR];
  uint32_t li[KPR];
#pragma unroll
  for (int e = 0; e < KPR; ++e) {
    ls[e] = e < kp ? (qvalid ? -INFINITY : INFINITY) : INFINITY;
    li[e] = NONE;
  }
  float th = ls[0];
  uint32_t my_inserts = 0;
  if (qs * 32 < nq) {  // the whole wave's queries lie beyond the block otherwise
    const float4* blk = reinterpret_cast<const float4*>(S + (size_t)qs * ncs * 1024 + l * 16);
    f32x16 nxt;
    auto fetch = [&](uint32_t cs) __attribute__((always_inline)) {
      const float4* src = blk + (size_t)cs * 256;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float4 f = src[v];
        nxt[4 * v] = f.x;
        nxt[4 * v + 1] = f.y;
        nxt[4 * v + 2] = f.z;
        nxt[4 * v + 3] = f.w;
      }
    };
    if (cs0 < cs1) fetch(cs0);
    for (uint32_t cs = cs0; cs < cs1; ++cs) {
      const f32x16 a = nxt;
      if (cs + 1 < cs1) fetch(cs + 1);
      const float m = max16(a);
      if (__builtin_amdgcn_ballot_w64(m > th) != 0 && m > th) {
        uint32_t mask = 0;
#pragma unroll
        for (int i = 15; i >= 0; --i) mask = (mask << 1) | (a[i] > th ? 1u : 0u);
        const uint32_t cbase = cs * 32 + 4 * h;
        while (mask) {
          const int i = __builtin_ctz(mask);
          mask &= mask - 1;
          const float v = pick16(a, i);
          const uint32_t cand = cbase + 8 * (i >> 2) + (i & 3);
          if (v > th && cand < c_end && cand != qg) {
            list_push<KPR>(ls, li, v, cand);
            th = ls[0];
            ++my_inserts;
          }
        }
      }
    }
  }
  if (inserts) {
    uint32_t tot = my_inserts;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (l == 0) atomicAdd(inserts, (unsigned long long)tot);
  }
  if (qvalid) {
    const size_t slot = ((size_t)qg * nsplit + split) * 2 + h;
    const size_t base = slot * (size_t)kp;
    uint32_t c = 0;
#pragma unroll
    for (int e = 0; e < KPR; ++e) {
      if (e < kp && li[e] != NONE) {
        cand_idx[base + c] = li[e];
        cand_sc[base + c] = ls[e];
        ++c;
      }
    }
    cand_cnt[slot] = c;
    cand_floor[slot] = th;
  }
}

// ------------------------------------------------------------------------------------------------
// Sorting helpers: bitonic sort of P (power of two) entries in LDS into "ranks_before" order,
// by NT threads of one workgroup.

template <int NT>
__device__ void bitonic_lds(double* __restrict__ sc, double* __restrict__ cs, uint32_t* __restrict__ id,
                            int P) {
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = tid_x(); e < P / 2; e += NT) {
        int a = 2 * e - (e & (stride - 1));
        int b = a + stride;
        bool up = (a & size) == 0;
        bool ab = ranks_before(sc[a], id[a], sc[b], id[b]);
        if (up != ab) {
          double ts = sc[a];
          sc[a] = sc[b];
          sc[b] = ts;
          if (cs) {
            double tc = cs[a];
            cs[a] = cs[b];
            cs[b] = tc;
          }
          uint32_t ti = id[a];
          id[a] = id[b];
          id[b] = ti;
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------------------------------------
// rescore_kernel: one 64-thread workgroup per query (as many waves in flight as the chip holds:
// the work is latency-bound).  Gathers the query's nsplit lists, rescores each candidate in f64
// over the original rows, sorts, applies the certificate and writes the top k (or queues the row
// for the fallback).

// The largest of the wave's err values into one of ERR_SLOTS maxima (err_max[workgroup % ERR_SLOTS],
// one atomic per wave at most; the host takes the largest slot).  Non-negative doubles order like
// their bit patterns read as unsigned integers.
constexpr uint32_t ERR_SLOTS = 256;
__device__ __forceinline__ void wave_err_max(double err, unsigned long long* __restrict__ err_slots) {
  unsigned long long* err_max = err_slots + (bid_x() & (ERR_SLOTS - 1));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) err = fmax(err, __shfl_xor(err, off));
  // only a wave that beats the value already in its slot (read first, without an atomic) takes
  // the atomic, so the updates stay few and spread over the slots
  const unsigned long long bits = (unsigned long long)__double_as_longlong(err);
  if ((tid_x() & 63) == 0 && err > 0.0 &&
      bits > __hip_atomic_load(err_max, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(err_max, bits);
}

// One query's rescoring (NT threads, a multiple of 64: all of them gather and rescore the entries,
// the first wave ranks them).  Failed rows are recorded as fb_tag in fb_rows.
// err_max collects the largest |list score - f64 score| over the row's admitted candidates: the
// measured counterpart of the error bound eps_list that the certificate relies on.  Scores are
// cosines, or with Tanimoto selection Tanimoto scores (ranked by them; the cosines of the k kept
// are computed again for the output).  Row j of the original rows starts at x + j * d; with V4, d
// is the stride of rows padded as dot64r<true> reads them.
template <int CMAX, bool V4, int NT = 64>
__device__ void rescore_row(double* __restrict__ s_sc, uint32_t* __restrict__ s_id, const float* __restrict__ x,
                            uint32_t d, const double* __restrict__ norms, uint32_t k, int kp, uint32_t nsplit,
                            uint32_t qq, uint32_t i, const uint32_t* __restrict__ cand_idx,
                            const float* __restrict__ cand_sc, const uint32_t* __restrict__ cand_cnt,
                            const float* __restrict__ cand_floor, int all_admitted, double eps_list, double eps32,
                            uint32_t select, uint32_t* __restrict__ idx, double* __restrict__ cos_out,
                            double* __restrict__ wts, uint32_t fb_tag, uint32_t* __restrict__ fb_rows,
                            double* __restrict__ fb_lo, uint32_t* __restrict__ fb_count,
                            uint32_t* __restrict__ short_count, unsigned long long* __restrict__ err_max) {
  const double ni = norms[i];
  const float* xi = x + (size_t)i * d;
  const int tot_slots = (int)nsplit * kp;
  const bool tani = select == COSTAL_SELECT_TANIMOTO;

  uint32_t total = 0;
  double floor_ = -INFINITY;
  for (uint32_t s = 0; s < nsplit; ++s) {
    total += cand_cnt[(size_t)qq * nsplit + s];
    floor_ = fmax(floor_, (double)cand_floor[(size_t)qq * nsplit + s]);
  }
  // the lists' entries, compacted (lists are disjoint, so the sort below fixes the order whatever
  // the order of arrival), then padded to a power of two for the sort
  __shared__ uint32_t s_fill;
  if (tid_x() == 0) s_fill = 0;
  __syncthreads();
  double err = 0.0;  // the largest |list score - f64 cosine| seen by this thread
  for (int e = tid_x(); e < tot_slots; e += NT) {
    const uint32_t s = (uint32_t)e / (uint32_t)kp, r = (uint32_t)e % (uint32_t)kp;
    const size_t slot = (size_t)qq * nsplit + s;
    if (r < cand_cnt[slot]) {
      const uint32_t j = cand_idx[slot * kp + r];
      const uint32_t pos = atomicAdd(&s_fill, 1u);
      const double c = score64(dot64r<V4>(xi, x + (size_t)j * d, d), ni, norms[j], tani);
      err = fmax(err, fabs((double)cand_sc[slot * kp + r] - c));
      s_sc[pos] = c;
      s_id[pos] = j;
    }
  }
  wave_err_max(err, err_max);
  __syncthreads();
  const uint32_t filled = s_fill;  // = total
  if (filled <= 64) {
    // one entry per thread: its rank is the number of entries ranking before it (the candidates
    // are distinct, so the ranks are a permutation), and it moves straight to that place
    const uint32_t e = tid_x();
    double sc_e = 0.0;
    uint32_t id_e = 0, rank = 0;
    if (e < filled) {
      sc_e = s_sc[e];
      id_e = s_id[e];
      for (uint32_t f = 0; f < filled; ++f) rank += ranks_before(s_sc[f], s_id[f], sc_e, id_e) ? 1u : 0u;
    }
    __syncthreads();
    if (e < filled) {
      s_sc[rank] = sc_e;
      s_id[rank] = id_e;
    }
    __syncthreads();
  } else {
    // More entries than threads (rescans with many splits): only the k best are needed, in order,
    // so k rounds of a wave-wide arg-max under the same order (each thread takes the best entry
    // it holds, the wave the best of those, which then leaves the pool) instead of a full sort
    // (k <= 64: round r's winner stays with thread r).  A bitonic sort of 1,024 entries by one
    // wave took 133 us per row.  With several waves each takes the best of its threads', and the
    // waves' bests meet in LDS (the order is total, so every thread picks the same winner).
    const uint32_t t = tid_x();
    const uint32_t rounds = k < filled ? k : filled;
    constexpr int NWR = NT / 64;
    __shared__ double s_wbs[NWR];
    __shared__ uint32_t s_wbi[NWR], s_wbe[NWR];
    double win_sc = -INFINITY;
    uint32_t win_id = NONE;
    for (uint32_t r = 0; r < rounds; ++r) {
      double bs = -INFINITY;
      uint32_t bi = NONE, be = NONE;
      for (uint32_t e = t; e < filled; e += NT) {
        const uint32_t id = s_id[e];
        const double sc = s_sc[e];
        if (id != NONE && (bi == NONE || ranks_before(sc, id, bs, bi))) bs = sc, bi = id, be = e;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double os = __shfl_xor(bs, off);
        const uint32_t oi = (uint32_t)__shfl_xor((int)bi, off), oe = (uint32_t)__shfl_xor((int)be, off);
        if (oi != NONE && (bi == NONE || ranks_before(os, oi, bs, bi))) bs = os, bi = oi, be = oe;
      }
      if constexpr (NWR > 1) {
        if ((t & 63) == 0) s_wbs[t >> 6] = bs, s_wbi[t >> 6] = bi, s_wbe[t >> 6] = be;
        __syncthreads();
#pragma unroll
        for (int v = 0; v < NWR; ++v) {
          const double os = s_wbs[v];
          const uint32_t oi = s_wbi[v], oe = s_wbe[v];
          if (oi != NONE && (bi == NONE || ranks_before(os, oi, bs, bi))) bs = os, bi = oi, be = oe;
        }
      }
      if (t == r) win_sc = bs, win_id = bi;
      if (t == 0) s_id[be] = NONE;
      __syncthreads();
    }
    if (t < rounds) {
      s_sc[t] = win_sc;
      s_id[t] = win_id;
    }
    __syncthreads();
  }

  bool ok;
  if (all_admitted) {
    ok = true;
  } else if (total < k) {
    ok = false;
  } else {
    ok = s_sc[k - 1] > floor_ + eps_list;
  }
  if (ok) {
    for (uint32_t t = tid_x(); t < k; t += NT) {
      const uint32_t j = s_id[t];
      const double c = tani ? cos64(dot64r<V4>(xi, x + (size_t)j * d, d), ni, norms[j]) : s_sc[t];
      idx[(size_t)i * k + t] = j;
      cos_out[(size_t)i * k + t] = c;
      if (wts) wts[(size_t)i * k + t] = tani_from_cos(ni, norms[j], c);
    }
  } else if (tid_x() == 0) {
    const uint32_t p = atomicAdd(fb_count, 1u);
    fb_rows[p] = fb_tag;
    fb_lo[p] = total >= k ? s_sc[k - 1] - eps32 : -INFINITY;
    if (total < k) atomicAdd(short_count, 1u);
  }
}

// Two queries per 64-thread workgroup, one per half-wave, for rows of at most 32 list entries:
// the same gather, f64 rescoring, rank placement, certificate and outputs as rescore_row, with
// twice the rows in flight (the work is latency-bound).  Both halves call this together with their
// own row (valid = false for a missing one), so every thread meets every barrier.
template <bool V4>
__device__ void rescore_half(double* __restrict__ s_sc, uint32_t* __restrict__ s_id, uint32_t* __restrict__ s_fill,
                             bool valid, uint32_t lane, const float* __restrict__ x, uint32_t d,
                             const double* __restrict__ norms, uint32_t k, int kp, uint32_t nsplit, uint32_t qq,
                             uint32_t i, const uint32_t* __restrict__ cand_idx, const float* __restrict__ cand_sc,
                             const uint32_t* __restrict__ cand_cnt, const float* __restrict__ cand_floor,
                             int all_admitted, double eps_list, double eps32, uint32_t select,
                             uint32_t* __restrict__ idx, double* __restrict__ cos_out, double* __restrict__ wts,
                             uint32_t fb_tag,
                             uint32_t* __restrict__ fb_rows, double* __restrict__ fb_lo, uint32_t* __restrict__ fb_count,
                             uint32_t* __restrict__ short_count, unsigned long long* __restrict__ err_max) {
  const int tot_slots = (int)nsplit * kp;  // <= 32
  const bool tani = select == COSTAL_SELECT_TANIMOTO;
  double ni = 0.0, floor_ = -INFINITY;
  uint32_t total = 0;
  if (valid) {
    ni = norms[i];
    for (uint32_t sp = 0; sp < nsplit; ++sp) {
      total += cand_cnt[(size_t)qq * nsplit + sp];
      floor_ = fmax(floor_, (double)cand_floor[(size_t)qq * nsplit + sp]);
    }
  }
  if (lane == 0) *s_fill = 0;
  __syncthreads();
  double err = 0.0;
  if (valid && (int)lane < tot_slots) {
    const uint32_t sp = lane / (uint32_t)kp, r = lane % (uint32_t)kp;
    const size_t slot = (size_t)qq * nsplit + sp;
    if (r < cand_cnt[slot]) {
      const uint32_t j = cand_idx[slot * kp + r];
      const uint32_t pos = atomicAdd(s_fill, 1u);
      const double c = score64(dot64r<V4>(x + (size_t)i * d, x + (size_t)j * d, d), ni, norms[j], tani);
      err = fmax(err, fabs((double)cand_sc[slot * kp + r] - c));
      s_sc[pos] = c;
      s_id[pos] = j;
    }
  }
  wave_err_max(err, err_max);
  __syncthreads();
  const uint32_t filled = *s_fill;
  double sc_e = 0.0;
  uint32_t id_e = 0, rank = 0;
  if (lane < filled) {
    sc_e = s_sc[lane];
    id_e = s_id[lane];
    for (uint32_t f = 0; f < filled; ++f) rank += ranks_before(s_sc[f], s_id[f], sc_e, id_e) ? 1u : 0u;
  }
  __syncthreads();
  if (lane < filled) {
    s_sc[rank] = sc_e;
    s_id[rank] = id_e;
  }
  __syncthreads();
  if (!valid) return;
  bool ok;
  if (all_admitted) ok = true;
  else if (total < k) ok = false;
  else ok = s_sc[k - 1] > floor_ + eps_list;
  if (ok) {
    for (uint32_t t = lane; t < k; t += 32) {
      const uint32_t j = s_id[t];
      const double c = tani ? cos64(dot64r<V4>(x + (size_t)i * d, x + (size_t)j * d, d), ni, norms[j]) : s_sc[t];
      idx[(size_t)i * k + t] = j;
      cos_out[(size_t)i * k + t] = c;
      if (wts) wts[(size_t)i * k + t] = tani_from_cos(ni, norms[j], c);
    }
  } else if (lane == 0) {
    const uint32_t pp = atomicAdd(fb_count, 1u);
    fb_rows[pp] = fb_tag;
    fb_lo[pp] = total >= k ? s_sc[k - 1] - eps32 : -INFINITY;
    if (total < k) atomicAdd(short_count, 1u);
  }
}

// The rows of a row list, or every row: in the locality order when perm is given (consecutive
// workgroups then rescore similar rows, whose candidates overlap, so their rows stay in L2: 10^6 x
// 40, k = 30: 3.54 -> 2.78 ms), lists by row.  With V4, x holds rows of d floats as dot64r<true>
// reads them (the input itself, or its padded copy).
template <bool V4>
__global__ __launch_bounds__(64) void rescore_pair_kernel(
    const float* __restrict__ x, const uint32_t* __restrict__ perm, uint32_t d, const double* __restrict__ norms,
    uint32_t k, int kp,
    uint32_t nsplit, const uint32_t* __restrict__ rowlist, uint32_t nq, const uint32_t* __restrict__ cand_idx,
    const float* __restrict__ cand_sc, const uint32_t* __restrict__ cand_cnt, const float* __restrict__ cand_floor,
    int all_admitted, double eps_list, double eps32, uint32_t select, uint32_t* __restrict__ idx,
    double* __restrict__ cos_out, double* __restrict__ wts, uint32_t* __restrict__ fb_rows, double* __restrict__ fb_lo,
    uint32_t* __restrict__ fb_count, uint32_t* __restrict__ short_count, unsigned long long* __restrict__ err_max) {
  __shared__ double s_sc[2][32];
  __shared__ uint32_t s_id[2][32];
  __shared__ uint32_t s_fill[2];
  const uint32_t half = tid_x() >> 5, lane = tid_x() & 31;
  const uint32_t qq = xcd_place(bid_x(), gdim_x()) * 2 + half;
  const bool valid = qq < nq;
  const uint32_t i = valid ? (rowlist ? rowlist[qq] : perm ? perm[qq] : qq) : 0;
  rescore_half<V4>(s_sc[half], s_id[half], &s_fill[half], valid, lane, x, d, norms, k, kp, nsplit, rowlist ? qq : i,
                   i,
               cand_idx,
               cand_sc, cand_cnt, cand_floor, all_admitted, eps_list, eps32, select, idx, cos_out, wts, i, fb_rows,
               fb_lo, fb_count, short_count, err_max);
}

// Threads per row of rescore_kernel: the few rows of the rescan over up to 96 splits gather and
// rank their entries with four waves, so that each thread's chain of dependent gathers and LDS
// reads stays short (two rows of 1,740 entries: 99 us with one wave, 50 us with four); rows of
// up to 1,024 entries come in thousands and keep one wave each (four: 62 -> 75 us for 2,876 rows).
constexpr int rescore_threads(int cmax) { return cmax >= 2048 ? 256 : 64; }

template <int CMAX, bool V4>
__global__ __launch_bounds__(rescore_threads(CMAX)) void rescore_kernel(
    const float* __restrict__ x, const uint32_t* __restrict__ perm, uint32_t d, const double* __restrict__ norms,
    uint32_t n,
    uint32_t k, int kp, uint32_t nsplit, const uint32_t* __restrict__ rowlist, uint32_t nq,
    const uint32_t* __restrict__ cand_idx, const float* __restrict__ cand_sc, const uint32_t* __restrict__ cand_cnt,
    const float* __restrict__ cand_floor, int all_admitted, double eps_list, double eps32, uint32_t select,
    uint32_t* __restrict__ idx, double* __restrict__ cos_out, double* __restrict__ wts,
    uint32_t* __restrict__ fb_rows, double* __restrict__ fb_lo, uint32_t* __restrict__ fb_count,
    uint32_t* __restrict__ short_count, unsigned long long* __restrict__ err_max) {
  __shared__ double s_sc[CMAX];
  __shared__ uint32_t s_id[CMAX];
  for (uint32_t qq = xcd_place(bid_x(), gdim_x()); qq < nq; qq += gdim_x()) {
    const uint32_t i = rowlist ? rowlist[qq] : perm ? perm[qq] : qq;
    rescore_row<CMAX, V4, rescore_threads(CMAX)>(s_sc, s_id, x, d, norms, k, kp, nsplit, rowlist ? qq : i, i, cand_idx,
                                                  cand_sc, cand_cnt, cand_floor, all_admitted, eps_list, eps32, select,
                                                  idx, cos_out, wts, i, fb_rows, fb_lo, fb_count, short_count, err_max);
    __syncthreads();  // the LDS arrays are refilled by the next row
  }
}

// ------------------------------------------------------------------------------------------------
// rowscan_kernel: exact top-k of whole rows.  One 256-thread workgroup per row (grid-stride over
// the row list).  Candidates stream in chunks of 256; survivors are appended to an LDS buffer that
// is merged into the running top-k (bitonic sort of list + buffer) when it nears capacity.
//   COARSE = true:  the survivor test is an f32 inner product of the normalized rows >= lo, with lo
//                   raised to (k-th exact score - eps) once k survivors are known (gather_row,
//                   knn.rs:723-743, with min_sim = s_k - eps).
//   COARSE = false: every candidate is scored exactly (knn_bruteforce_select, knn.rs:884-924),
//                   for either selection metric.

template <int CAP>
struct RowscanSmem {
  double sc[CAP];
  double cs[CAP];
  uint32_t id[CAP];
  int L, B, full;
  double lo, ksc;
  uint32_t kid;
};

// The exact top-k of row i (256 threads); lo0 is the initial survivor bound (-inf: none).
template <int CAP, bool COARSE>
__device__ void rowscan_row(RowscanSmem<CAP>& S, const float* __restrict__ x, uint32_t d,
                            const float* __restrict__ xs32, uint32_t dp32, const double* __restrict__ norms,
                            uint32_t n, uint32_t k, uint32_t select, uint32_t i, double lo0, double eps32,
                            uint32_t* __restrict__ idx, double* __restrict__ cos_out, double* __restrict__ wts) {
  const int tid = tid_x();
  if (tid == 0) {
    S.L = 0;
    S.B = 0;
    S.full = 0;
    S.lo = lo0;
    S.ksc = -INFINITY;
    S.kid = NONE;
  }
  __syncthreads();
  const double ni = norms[i];
  const float* xi = x + (size_t)i * d;

  auto merge = [&]() {
    const int tot = S.L + S.B;
    int P = 2;
    while (P < tot) P <<= 1;
    for (int e = tot + tid; e < P; e += 256) {
      S.sc[e] = -INFINITY;
      S.cs[e] = 0.0;
      S.id[e] = NONE;
    }
    __syncthreads();
    bitonic_lds<256>(S.sc, S.cs, S.id, P);
    if (tid == 0) {
      S.L = tot < (int)k ? tot : (int)k;
      S.B = 0;
      if (S.L == (int)k) {
        S.full = 1;
        S.ksc = S.sc[k - 1];
        S.kid = S.id[k - 1];
        if (COARSE) S.lo = fmax(S.lo, S.ksc - eps32);
      }
    }
    __syncthreads();
  };

  for (uint32_t base = 0; base < n; base += 256) {
    const uint32_t j = base + tid;
    if (j < n && j != i) {
      bool pass = true;
      if (COARSE) {
        // the f32 score of the normalized rows (for Tanimoto selection, of their cosine and the
        // raw norms in f32, as the fused scan scores it)
        float c32 = dot32(xs32 + (size_t)i * dp32, xs32 + (size_t)j * dp32, dp32);
        if (select == COSTAL_SELECT_TANIMOTO) {
          const float a = (float)ni;
          c32 = tani32(c32, (float)norms[j], a > 0.0f ? __builtin_amdgcn_rcpf(a) : 0.0f);
        }
        pass = (double)c32 >= S.lo;
      }
      if (pass) {
        const double nj = norms[j];
        const double c = cos64(dot64(xi, x + (size_t)j * d, d), ni, nj);
        const double sc = select == COSTAL_SELECT_TANIMOTO ? tani_from_cos(ni, nj, c) : c;
        if (!S.full || ranks_before(sc, j, S.ksc, S.kid)) {
          const int p = atomicAdd(&S.B, 1);
          const int slot = S.L + p;
          S.sc[slot] = sc;
          S.cs[slot] = c;
          S.id[slot] = j;
        }
      }
    }
    __syncthreads();
    const bool full_buffer = S.L + S.B > CAP - 256;
    __syncthreads();
    if (full_buffer) merge();
  }
  merge();
  for (uint32_t t = tid; t < k; t += 256) {
    const uint32_t j = S.id[t];
    idx[(size_t)i * k + t] = j;
    cos_out[(size_t)i * k + t] = S.cs[t];
    if (wts) wts[(size_t)i * k + t] = j < n ? tani_from_cos(ni, norms[j], S.cs[t]) : 0.0;
  }
  __syncthreads();
}

template <int CAP, bool COARSE>
__global__ __launch_bounds__(256) void rowscan_kernel(
    const float* __restrict__ x, uint32_t d, const float* __restrict__ xs32, uint32_t dp32,
    const double* __restrict__ norms, uint32_t n, uint32_t k, uint32_t select,
    const uint32_t* __restrict__ rows, const uint32_t* __restrict__ nrows_dev, uint32_t nrows_host,
    const double* __restrict__ lo_arr, double eps32, uint32_t* __restrict__ idx,
    double* __restrict__ cos_out, double* __restrict__ wts) {
  __shared__ RowscanSmem<CAP> S;
  const uint32_t nrows = nrows_dev ? *nrows_dev : nrows_host;
  for (uint32_t r = bid_x(); r < nrows; r += gdim_x()) {
    const uint32_t i = rows ? rows[r] : r;
    rowscan_row<CAP, COARSE>(S, x, d, xs32, dp32, norms, n, k, select, i, lo_arr ? lo_arr[r] : -INFINITY, eps32,
                             idx, cos_out, wts);
  }
}

// ------------------------------------------------------------------------------------------------
// Batched samples: many independent kNN problems (uneven n, d) in one pass of each kernel.  The
// working copies of all samples are stacked (each sample padded to 512 rows, all rows padded to the
// batch's widest feature count, zeros being exact), and one persistent scan kernel pulls
// (sample, query block, candidate split) work items from a queue ordered by decreasing cost.


// the sample a row belongs to, by binary search in an ascending table of first rows
__device__ __forceinline__ uint32_t find_sample(const uint32_t* __restrict__ first, uint32_t nsamples, uint32_t g) {
  uint32_t lo = 0, hi = nsamples;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (first[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// counters[0, none) = all ones, counters[none, count) = 0 (a kernel, so that a call needs no host
// copy and no memset node)
__global__ void counters_init_kernel(uint32_t* __restrict__ counters, uint32_t count, uint32_t none) {
  for (uint32_t e = tid_x(); e < count; e += bdim_x()) counters[e] = e < none ? NONE : 0u;
}

// Graph-safe counter reset (no memset nodes): counters[0..1] = first bad (sample, row) as u64
// (all ones = none), everything else zero.
__global__ void batch_init_kernel(uint32_t* __restrict__ counters, uint32_t count) {
  for (uint32_t e = tid_x(); e < count; e += bdim_x()) counters[e] = e < 2 ? NONE : 0u;
}

__global__ void prep_batch_kernel(const BatchSample* __restrict__ samples, const uint32_t* __restrict__ pad_first,
                                  uint32_t nsamples, uint32_t total_pad, uint32_t dp32, uint32_t dp16, int split16,
                                  float* __restrict__ xs32, _Float16* __restrict__ xs16,
                                  unsigned long long* __restrict__ bad) {
  const uint32_t g = bid_x() * bdim_x() + tid_x();
  if (g >= total_pad) return;
  const uint32_t sidx = find_sample(pad_first, nsamples, g);
  const BatchSample& S = samples[sidx];
  const uint32_t r = g - S.row_pad;
  const uint32_t s16 = split16 ? 2 * dp16 : dp16;
  // 16-byte stores throughout (dp32 is a multiple of 8 and dp16 of 16, so rows stay aligned)
  if (r >= S.n) {
    for (uint32_t t = 0; t < dp32; t += 4) *reinterpret_cast<f32x4*>(xs32 + (size_t)g * dp32 + t) = f32x4{};
    if (xs16)
      for (uint32_t t = 0; t < s16; t += 8) *reinterpret_cast<f16x8*>(xs16 + (size_t)g * s16 + t) = f16x8{};
    return;
  }
  // the row's bits (batched samples have d <= 64) in registers, read once (the non-finite test
  // reads them as integers): 16-byte loads when the row is aligned to them, the feature loops
  // unrolled to 64 so every index is static
  constexpr int DMAX = 64;
  const uint32_t d = S.d;
  const uint32_t* row = reinterpret_cast<const uint32_t*>(S.x + (size_t)r * d);
  uint32_t rb[DMAX];
  if ((d & 3) == 0 && (reinterpret_cast<uintptr_t>(row) & 15) == 0) {
#pragma unroll
    for (int t = 0; t < DMAX; t += 4) {
      uint4 q = {0u, 0u, 0u, 0u};
      if ((uint32_t)t < d) q = *reinterpret_cast<const uint4*>(row + t);
      rb[t] = q.x;
      rb[t + 1] = q.y;
      rb[t + 2] = q.z;
      rb[t + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int t = 0; t < DMAX; ++t) rb[t] = (uint32_t)t < d ? row[t] : 0u;
  }
  double sum = 0.0;
  bool badrow = false;
#pragma unroll
  for (int t = 0; t < DMAX; ++t) {
    if ((uint32_t)t < d) {
      badrow |= (rb[t] & 0x7f800000u) == 0x7f800000u;
      const float v = __uint_as_float(rb[t]);
      sum += (double)v * (double)v;
    }
  }
  const double nrm = sqrt(sum);
  S.norms[r] = nrm;
  if (badrow) atomicMin(bad, ((unsigned long long)sidx << 32) | r);
  float nv[DMAX];
#pragma unroll
  for (int t = 0; t < DMAX; ++t)
    nv[t] = ((uint32_t)t < d && nrm > 0.0) ? (float)((double)__uint_as_float(rb[t]) / nrm) : 0.0f;
#pragma unroll
  for (int t = 0; t < DMAX; t += 4)
    if ((uint32_t)t < dp32) *reinterpret_cast<f32x4*>(xs32 + (size_t)g * dp32 + t) = f32x4{nv[t], nv[t + 1], nv[t + 2], nv[t + 3]};
  if (xs16) {
#pragma unroll
    for (int t = 0; t < DMAX; t += 8) {
      if ((uint32_t)t < dp16) {
        f16x8 hi, lo;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          hi[c] = (_Float16)nv[t + c];
          lo[c] = (_Float16)(nv[t + c] - (float)hi[c]);
        }
        *reinterpret_cast<f16x8*>(xs16 + (size_t)g * s16 + t) = hi;
        if (split16) *reinterpret_cast<f16x8*>(xs16 + (size_t)g * s16 + dp16 + t) = lo;
      }
    }
  }
}


template <int CMAX>
__global__ __launch_bounds__(64) void rescore_batch_kernel(
    const BatchSample* __restrict__ samples, const uint32_t* __restrict__ row_first, uint32_t nsamples,
    uint32_t total_rows, uint32_t k, double eps_list, double eps32, const uint32_t* __restrict__ cand_idx,
    const float* __restrict__ cand_sc, const uint32_t* __restrict__ cand_cnt, const float* __restrict__ cand_floor,
    uint32_t* __restrict__ fb_rows, double* __restrict__ fb_lo, uint32_t* __restrict__ fb_count,
    uint32_t* __restrict__ short_count, unsigned long long* __restrict__ err_max,
    const unsigned long long* __restrict__ bad) {
  __shared__ double s_sc[CMAX];
  __shared__ uint32_t s_id[CMAX];
  const uint32_t g = bid_x();
  if (g >= total_rows || *bad != ~0ull) return;
  const uint32_t sidx = find_sample(row_first, nsamples, g);
  const BatchSample& S = samples[sidx];
  const uint32_t i = g - S.row_base;
  const uint32_t nlists = 2 * S.nsplit;
  const int all_admitted = (S.nsplit == 1 && (uint32_t)S.kp >= S.n - 1) ? 1 : 0;
  rescore_row<CMAX, false>(s_sc, s_id, S.x, S.d, S.norms, k, S.kp, nlists, i, i, cand_idx + S.entry_base,
                    cand_sc + S.entry_base, cand_cnt + S.slot_base, cand_floor + S.slot_base, all_admitted, eps_list,
                    eps32, COSTAL_SELECT_COSINE, S.idx, S.cos, S.wts, g, fb_rows, fb_lo, fb_count, short_count,
                    err_max);
}

// rescore_batch_kernel for batches whose rows hold at most 32 list entries: two rows per wave
__global__ __launch_bounds__(64) void rescore_batch_pair_kernel(
    const BatchSample* __restrict__ samples, const uint32_t* __restrict__ row_first, uint32_t nsamples,
    uint32_t total_rows, uint32_t k, double eps_list, double eps32, const uint32_t* __restrict__ cand_idx,
    const float* __restrict__ cand_sc, const uint32_t* __restrict__ cand_cnt, const float* __restrict__ cand_floor,
    uint32_t* __restrict__ fb_rows, double* __restrict__ fb_lo, uint32_t* __restrict__ fb_count,
    uint32_t* __restrict__ short_count, unsigned long long* __restrict__ err_max,
    const unsigned long long* __restrict__ bad) {
  __shared__ double s_sc[2][32];
  __shared__ uint32_t s_id[2][32];
  __shared__ uint32_t s_fill[2];
  if (*bad != ~0ull) return;  // the whole workgroup leaves together
  const uint32_t half = tid_x() >> 5, lane = tid_x() & 31;
  const uint32_t g = bid_x() * 2 + half;
  const bool valid = g < total_rows;
  const BatchSample& S = samples[valid ? find_sample(row_first, nsamples, g) : 0];
  const uint32_t i = valid ? g - S.row_base : 0;
  const int all_admitted = (S.nsplit == 1 && (uint32_t)S.kp >= S.n - 1) ? 1 : 0;
  rescore_half<false>(s_sc[half], s_id[half], &s_fill[half], valid, lane, S.x, S.d, S.norms, k, S.kp, 2 * S.nsplit,
                      i, i,
               cand_idx + S.entry_base, cand_sc + S.entry_base, cand_cnt + S.slot_base, cand_floor + S.slot_base,
               all_admitted, eps_list, eps32, COSTAL_SELECT_COSINE, S.idx, S.cos, S.wts, g, fb_rows, fb_lo, fb_count,
               short_count, err_max);
}

template <int CAP>
__global__ __launch_bounds__(256) void rowscan_batch_kernel(
    const BatchSample* __restrict__ samples, const uint32_t* __restrict__ row_first, uint32_t nsamples,
    const float* __restrict__ xs32, uint32_t dp32, uint32_t k, const uint32_t* __restrict__ rows,
    const uint32_t* __restrict__ nrows_dev, const double* __restrict__ lo_arr, double eps32,
    const unsigned long long* __restrict__ bad) {
  __shared__ RowscanSmem<CAP> Sm;
  const uint32_t nrows = *bad != ~0ull ? 0u : *nrows_dev;
  for (uint32_t r = bid_x(); r < nrows; r += gdim_x()) {
    const uint32_t g = rows[r];
    const uint32_t sidx = find_sample(row_first, nsamples, g);
    const BatchSample& S = samples[sidx];
    rowscan_row<CAP, true>(Sm, S.x, S.d, xs32 + (size_t)S.row_pad * dp32, dp32, S.norms, S.n, k,
                           COSTAL_SELECT_COSINE, g - S.row_base, lo_arr[r], eps32, S.idx, S.cos, S.wts);
  }
}

// ------------------------------------------------------------------------------------------------
// Tanimoto selection inputs of the fused scan.

// The largest row norm (ext[0..1], f64 bits) and the smallest positive one (ext[2..3], the
// complement of its bits, so that both are maxima from zero): a grid-stride pass, one pair of
// atomics per workgroup (per wave they queued on the one address: 78 us on 396,460 rows).
__global__ __launch_bounds__(256) void norm_extremes_kernel(const double* __restrict__ norms, uint32_t n,
                                                            unsigned long long* __restrict__ ext) {
  __shared__ unsigned long long s_hi[4], s_lo[4];
  unsigned long long hi = 0, lo = 0;
  for (uint32_t i = bid_x() * 256 + tid_x(); i < n; i += gdim_x() * 256) {
    const double v = norms[i];
    hi = std::max(hi, (unsigned long long)__double_as_longlong(v));
    if (v > 0.0) lo = std::max(lo, ~(unsigned long long)__double_as_longlong(v));
  }
  for (int off = 32; off > 0; off >>= 1) {
    hi = std::max(hi, (unsigned long long)__shfl_xor((long long)hi, off));
    lo = std::max(lo, (unsigned long long)__shfl_xor((long long)lo, off));
  }
  if ((tid_x() & 63) == 0) s_hi[tid_x() >> 6] = hi, s_lo[tid_x() >> 6] = lo;
  __syncthreads();
  if (tid_x() == 0) {
    hi = std::max(std::max(s_hi[0], s_hi[1]), std::max(s_hi[2], s_hi[3]));
    lo = std::max(std::max(s_lo[0], s_lo[1]), std::max(s_lo[2], s_lo[3]));
    if (hi) atomicMax(&ext[0], hi);
    if (lo) atomicMax(&ext[1], lo);
  }
}

// The raw norms of the working rows times `scale` (a power of two that keeps them in f32 range;
// Tanimoto scores depend on norm ratios only) in f32, rows past n zero, and per block of `half`
// working rows the smallest and largest scaled norm rounded outward to f32 ((0, 0) past n): one
// wave per block.
__global__ __launch_bounds__(64) void tani_norms_kernel(const double* __restrict__ norms,
                                                        const uint32_t* __restrict__ perm, uint32_t n, uint32_t half,
                                                        double scale, float* __restrict__ tnorm,
                                                        float2* __restrict__ trange) {
  const uint32_t b = bid_x(), l = tid_x();
  double lo = INFINITY, hi = -INFINITY;
  for (uint32_t r = l; r < half; r += 64) {
    const uint32_t w = b * half + r;
    float f = 0.0f;
    if (w < n) {
      const double v = norms[perm ? perm[w] : w] * scale;
      lo = fmin(lo, v);
      hi = fmax(hi, v);
      f = (float)v;
    }
    tnorm[w] = f;
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off));
    hi = fmax(hi, __shfl_xor(hi, off));
  }
  if (l == 0) {
    float2 r = make_float2(0.0f, 0.0f);
    if (hi >= lo) {
      r.x = (float)lo;
      if ((double)r.x > lo) r.x = nextafterf(r.x, -INFINITY);
      r.y = (float)hi;
      if ((double)r.y < hi) r.y = nextafterf(r.y, INFINITY);
    }
    trange[b] = r;
  }
}

// ------------------------------------------------------------------------------------------------
// weight::reweight (weight.rs:41-60): tanimoto / cosine.  Jaccard runs on the host.

// Admission thresholds of the rescan, by row: fb_lo - eps_list rounded down to f32, or -inf for a
// row with fewer than k candidates (fb_lo = -inf).
__global__ void rescan_thresholds_kernel(const uint32_t* __restrict__ rows, const double* __restrict__ lo, uint32_t nrows,
                                         double eps_list, float* __restrict__ th) {
  const uint32_t q = bid_x() * bdim_x() + tid_x();
  if (q >= nrows) return;
  float t = -INFINITY;
  if (lo[q] > -INFINITY) {
    const double want = lo[q] - eps_list;
    t = (float)want;
    if ((double)t > want) {  // one f32 step down (want is finite: |fb_lo| <= 1 + eps32)
      const uint32_t b = __float_as_uint(t);
      t = t > 0.0f ? __uint_as_float(b - 1) : t < 0.0f ? __uint_as_float(b + 1) : -__uint_as_float(1u);
    }
  }
  th[rows[q]] = t;
}

// Second-chance thresholds of the rows left short by the learned ones (lo = -inf; the others keep
// theirs): th = 1 - e - margin with e = exp(u + offset), so e * exp(offset2 - offset) =
// exp(u + offset2).  Any threshold keeps the result exact (a row short again is rescanned without
// one); this one only has to be low.
__global__ void second_chance_kernel(const uint32_t* __restrict__ rows, const double* __restrict__ lo, uint32_t nrows,
                                     const float* __restrict__ th, float margin, float factor,
                                     float* __restrict__ th2) {
  const uint32_t q = bid_x() * bdim_x() + tid_x();
  if (q >= nrows || lo[q] > -INFINITY) return;
  const uint32_t r = rows[q];
  th2[r] = 1.0f - (1.0f - th[r] - margin) * factor - margin;
}

// The failures of a rescan split by kind: rows with a k-th score (lo finite) to (bounded,
// bounded_lo), rows with fewer than k candidates to unbounded; counts in *nb, *nu.
__global__ void split_failures_kernel(const uint32_t* __restrict__ rows, const double* __restrict__ lo, uint32_t nrows,
                                      uint32_t* __restrict__ bounded, double* __restrict__ bounded_lo,
                                      uint32_t* __restrict__ nb, uint32_t* __restrict__ unbounded,
                                      uint32_t* __restrict__ nu) {
  const uint32_t q = bid_x() * bdim_x() + tid_x();
  if (q >= nrows) return;
  if (lo[q] > -INFINITY) {
    const uint32_t p = atomicAdd(nb, 1u);
    bounded[p] = rows[q];
    bounded_lo[p] = lo[q];
  } else {
    unbounded[atomicAdd(nu, 1u)] = rows[q];
  }
}

__global__ void reweight_kernel(uint64_t n, uint32_t k, const uint32_t* __restrict__ idx,
                                const double* __restrict__ cos, const double* __restrict__ norms,
                                uint32_t weight, double* __restrict__ out,
                                uint32_t* __restrict__ bad) {
  const uint64_t e = (uint64_t)bid_x() * bdim_x() + tid_x();
  if (e >= n * k) return;
  const uint32_t j = idx[e];
  if (j >= n) {
    atomicMin(bad, (uint32_t)(e < 0xffffffffull ? e : 0xfffffffeull));
    return;
  }
  out[e] = weight == COSTAL_WEIGHT_TANIMOTO ? tani_from_cos(norms[e / k], norms[j], cos[e]) : cos[e];
}

// ------------------------------------------------------------------------------------------------
// Self-check of the two matrix-core properties the split-FP16 error bound relies on: how an MFMA
// rounds its FP32 accumulation, and whether FP16 subnormal inputs are kept.  One wave; only lane 0
// holds non-zero operands, so every probe's result is D[0][0] = sum_j a_j b_j + c (lane 0's
// halves are matching k indices in the A and B layouts), held by lane 0 in accumulator 0.
//   out[0] c = 1, one product 2^-24 + 2^-30 (just above half an ulp of 1): 1 + 2^-23 when the sum
//          is rounded to nearest, 1 when it is truncated
//   out[1] c = 1, three products 2^-25 (a quarter ulp each): 1 + 2^-23 if the sum is rounded once,
//          1 if each addition is rounded in turn (both round to nearest)
//   out[2] c = 1, one product -(2^-24 + 2^-30): 1 - 2^-24 to nearest, 1 - 2^-23 truncated
//   out[3] c = 0, 2^-24 (the smallest FP16 subnormal) x 1: 2^-24 if subnormals are kept, else 0
//   out[4] c = 0, 2^-20 (subnormal) x 2^-4: 2^-24 if kept, else 0
//   out[5] FP32 MFMA (32x32x2), c = 1, one product 2^-24 + 2^-30: 1 + 2^-23 to nearest
//   out[6] FP32 MFMA, c = 1, one product -(2^-24 + 2^-30): 1 - 2^-24 to nearest
//   out[7] as out[3], the subnormal built from its bits (0x0001) instead of by conversion
__global__ __launch_bounds__(64) void mfma_probe_kernel(float* __restrict__ out) {
  const bool l0 = tid_x() == 0;
  auto run16 = [&](float a0, float a1, float a2, float b0, float b1, float b2, float c) -> float {
    f16x8 A = {}, B = {};
    f32x16 C = {};
    if (l0) {
      A[0] = (_Float16)a0, A[1] = (_Float16)a1, A[2] = (_Float16)a2;
      B[0] = (_Float16)b0, B[1] = (_Float16)b1, B[2] = (_Float16)b2;
      C[0] = c;
    }
    const f32x16 D = __builtin_amdgcn_mfma_f32_32x32x16_f16(A, B, C, 0, 0, 0);
    return D[0];
  };
  auto run32 = [&](float a, float b, float c) -> float {
    f32x16 C = {};
    if (l0) C[0] = c;
    const f32x16 D = __builtin_amdgcn_mfma_f32_32x32x2f32(l0 ? a : 0.0f, l0 ? b : 0.0f, C, 0, 0, 0);
    return D[0];
  };
  float r[8];
  r[0] = run16(0x1.04p-12f, 0.0f, 0.0f, 0x1p-12f, 0.0f, 0.0f, 1.0f);
  r[1] = run16(0x1p-12f, 0x1p-12f, 0x1p-12f, 0x1p-13f, 0x1p-13f, 0x1p-13f, 1.0f);
  r[2] = run16(-0x1.04p-12f, 0.0f, 0.0f, 0x1p-12f, 0.0f, 0.0f, 1.0f);
  r[3] = run16(0x1p-24f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f);
  r[4] = run16(0x1p-20f, 0.0f, 0.0f, 0x1p-4f, 0.0f, 0.0f, 0.0f);
  r[5] = run32(0x1.04p-12f, 0x1p-12f, 1.0f);
  r[6] = run32(-0x1.04p-12f, 0x1p-12f, 1.0f);
  {
    f16x8 A = {}, B = {};
    f32x16 C = {};
    if (l0) {
      A[0] = __builtin_bit_cast(_Float16, (uint16_t)0x0001u);
      B[0] = (_Float16)1.0f;
    }
    r[7] = __builtin_amdgcn_mfma_f32_32x32x16_f16(A, B, C, 0, 0, 0)[0];
  }
  if (l0)
    for (int e = 0; e < 8; ++e) out[e] = r[e];
}

// ------------------------------------------------------------------------------------------------
// Host side

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(COSTAL_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

struct Device {
  std::mutex mu;
  bool init = false;
  int id = 0;
  int cus = 0;
  hipStream_t stream = nullptr;
  void* ws = nullptr;
  size_t ws_cap = 0;
  // device copies of the host entry points' arguments (grow-only, kept between calls)
  void* io = nullptr;
  size_t io_cap = 0;
  // this device's copy of the static scratch (hipGetSymbolAddress at set-up), or null with the
  // lookup's error
  void* scratch = nullptr;
  hipError_t scratch_err = hipSuccess;
  // call start/end and per-kernel brackets: prep, coarse, rescore, fallback
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evk[8] = {};
  // pinned host memory for the small readbacks the host branches on (counters), made by the
  // allocating entry points; null until then, and for a caller that never allocates
  void* hbuf = nullptr;
  // a second stream (default priority) for work that overlaps the call's own: the learned
  // thresholds run there beside the locality order; fork and join events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

// One state per device, selected by the device current at the call (a process may drive several
// GPUs, one thread each).
constexpr int MAX_DEVICES = 16;
Device g_devs[MAX_DEVICES];

Device* call_device() {
  int id = 0;
  hipError_t e = hipGetDevice(&id);
  if (e != hipSuccess) {
    fail(COSTAL_E_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    return nullptr;
  }
  if (id < 0 || id >= MAX_DEVICES) {
    fail(COSTAL_E_DEVICE, "device %d is beyond the %d devices this library serves", id, MAX_DEVICES);
    return nullptr;
  }
  g_devs[id].id = id;
  return &g_devs[id];
}

int device_init(Device& D) {
  if (D.init) return 0;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(COSTAL_E_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, D.id));
  D.cus = prop.multiProcessorCount;
  HIP_TRY(hipStreamCreateWithFlags(&D.stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&D.ev0));
  HIP_TRY(hipEventCreate(&D.ev1));
  HIP_TRY(hipStreamCreateWithFlags(&D.side, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&D.ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&D.ev_join, hipEventDisableTiming));
  for (auto& e : D.evk) HIP_TRY(hipEventCreate(&e));
  // the device's own address of the static scratch (the device is current here)
  D.scratch_err = hipGetSymbolAddress(&D.scratch, HIP_SYMBOL(scratch));
  if (D.scratch_err != hipSuccess) {
    D.scratch = nullptr;
    (void)hipGetLastError();
  }
  D.init = true;
  return 0;
}

constexpr size_t HBUF_BYTES = 4096;

// Copies bytes at src (device) to dst (host) once the work queued on s before it is done, and
// waits for it.  Through the device's pinned buffer when it has one: a copy into pageable memory
// is staged by the runtime, which adds to each of the call's host round trips.
int readback(Device& D, void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (D.hbuf && bytes <= HBUF_BYTES) {
    HIP_TRY(hipMemcpyAsync(D.hbuf, src, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    memcpy(dst, D.hbuf, bytes);
  } else {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  return 0;
}

int ws_reserve(Device& D, size_t bytes) {
  if (!D.hbuf && hipHostMalloc(&D.hbuf, HBUF_BYTES, hipHostMallocDefault) != hipSuccess) {
    D.hbuf = nullptr;  // readbacks go to pageable memory
    (void)hipGetLastError();
  }
  if (bytes <= D.ws_cap) return 0;
  if (D.ws) {
    HIP_TRY(hipStreamSynchronize(D.stream));
    HIP_TRY(hipFree(D.ws));
    D.ws = nullptr;
    D.ws_cap = 0;
  }
  size_t cap = bytes + bytes / 8 + (1u << 20);
  HIP_TRY(hipMalloc(&D.ws, cap));
  D.ws_cap = cap;
  return 0;
}

// The host entry points' device buffers: grow-only like the workspace, so a call of the same or a
// smaller size allocates nothing.
int io_reserve(Device& D, size_t bytes) {
  if (bytes <= D.io_cap) return 0;
  if (D.io) {
    HIP_TRY(hipStreamSynchronize(D.stream));
    HIP_TRY(hipFree(D.io));
    D.io = nullptr;
    D.io_cap = 0;
  }
  const size_t cap = bytes + bytes / 8 + (1u << 20);
  HIP_TRY(hipMalloc(&D.io, cap));
  D.io_cap = cap;
  return 0;
}

// Counter block of a call: 64 words of counters, then ERR_SLOTS u64 maxima of the measured error.
constexpr uint32_t COUNTER_WORDS = 64 + 2 * ERR_SLOTS;

// the largest of the ERR_SLOTS error maxima (f64 bit patterns, two words each)
double slots_max(const uint32_t* w) {
  double m = 0.0;
  for (uint32_t e = 0; e < ERR_SLOTS; ++e) {
    const uint64_t b = (uint64_t)w[2 * e] | ((uint64_t)w[2 * e + 1] << 32);
    double v;
    memcpy(&v, &b, sizeof v);
    m = std::max(m, v);
  }
  return m;
}

// bump allocator over the workspace
struct Carve {
  uint8_t* base;
  size_t off = 0;
  explicit Carve(void* b) : base((uint8_t*)b) {}
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T);
    return p;
  }
};

inline uint32_t round_up(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

// The fused path handles d <= 64 (FP32) / 128 (FP16) and K' <= 64; everything else runs the
// exact row scan.
constexpr uint32_t FUSED_KP_MAX = 64;
constexpr uint32_t RESCORE_CMAX = 2048;  // list entries one row's rescoring holds (passes over many rows)
// ... and for the few rows of the last rescan (48 KB of LDS per row: a rescoring workgroup each)
constexpr uint32_t RESCORE_CMAX_FEW = 4096;
constexpr uint32_t ROWSCAN_K_MAX = 2048;

struct Plan {
  bool fused;
  bool tani;        // Tanimoto selection (the fused scan of split FP16 rows with Tanimoto lists)
  bool wide;        // wide rows: score blocks (wide_gemm_kernel) + wide_scan_kernel
  uint32_t wide_qb; // query rows per score block (wide)
  int prec;
  uint32_t J;       // 16-byte chunks per lane half per row
  uint32_t kp;      // K'
  int KP;           // list capacity (template)
  int NW;           // waves per workgroup
  uint32_t QB, TC;  // queries per workgroup, candidates per tile
  uint32_t nsplit, cols_per_split;
  uint32_t dp32, dp16, n_pad;
  bool prune;       // first pass in the locality order with tile pruning (order.h)
  costal::OrderShape order;
};

using costal::CoarseFn;

CoarseFn coarse_select(int prec, uint32_t J, int KP, bool tani = false) {
  using namespace costal;
  if (tani) return KP == 16 ? scan_tani_fn_k16(J) : KP == 32 ? scan_tani_fn_k32(J) : scan_tani_fn_k64(J);
  if (prec == PREC_F32) return KP == 16 ? scan_fn_p0_k16(J) : KP == 32 ? scan_fn_p0_k32(J) : scan_fn_p0_k64(J);
  if (prec == PREC_F16) return KP == 16 ? scan_fn_p1_k16(J) : KP == 32 ? scan_fn_p1_k32(J) : scan_fn_p1_k64(J);
  return KP == 16 ? scan_fn_p2_k16(J) : KP == 32 ? scan_fn_p2_k32(J) : scan_fn_p2_k64(J);
}

// FP16 candidate error bound.  x16 = fp16(fp32(x/|x|)): |x16 - x| <= U|x| + eta per component with
// U = 2^-11 (1 + 2^-24) + 2^-24 and eta = 2^-25 (half the FP16 subnormal spacing).  With |x| = 1,
// sum |x16 y16 - x y| <= (U + eta sqrt d)(2 + U + eta sqrt d); the FP32 accumulation of the exact
// FP16 products adds gamma_d (1 + U + eta sqrt d)^2, gamma_m = m u / (1 - m u) with u = 2^-24: the
// matrix core rounds its FP32 accumulation to nearest (costal_mfma_probe shows it, and that FP16
// subnormal inputs are kept), and gamma_m bounds any order of m rounded additions.  eps_pre also
// covers the f32 re-score error ((d+4) 2^-23) and the rounding of (threshold - eps_pre) in f32.
double fp16_eps(uint32_t d) {
  const double u32 = std::ldexp(1.0, -24), uacc = u32;
  const double U = std::ldexp(1.0, -11) * (1.0 + u32) + u32;
  const double e = std::ldexp(1.0, -25) * std::sqrt((double)d);
  const double gam = d * uacc / (1.0 - d * uacc);
  return (U + e) * (2.0 + U + e) + gam * (1.0 + U + e) * (1.0 + U + e);
}

double fp32_eps(uint32_t d) { return ((double)d + 4.0) * (double)std::ldexp(1.0, -23); }

// Bound on |f32 Tanimoto score - f64 Tanimoto score| for list scores computed from cosines within e
// of the f64 cosine: the slope of T = c / (s - c) is s / (s - c)^2 <= 2 / (1 - e)^2 for s >= 2 and
// c <= 1 + e, and the f32 evaluation (tani32: about 16 roundings of a value <= 1 + 3e) adds < 2^-19.
double tani_eps(double e) { return 2.0 * e / ((1.0 - e) * (1.0 - e)) + std::ldexp(1.0, -19); }

// Split-FP16 bound.  v = f32(x/|x|) (relative error 2^-24), hi = fp16(v), lo = fp16(v - hi): the
// represented value hi + lo is within a|x| + eta of x per component, a = 2^-22 (1 + 2^-24) + 2^-24.
// The MFMA sum hi.hi + hi.lo + lo.hi differs from (hi+lo).(hi+lo) by the dropped lo.lo
// (<= (2^-11 + eta sqrt d)^2), the represented rows differ from the unit rows by
// (a + eta sqrt d)(2 + a + eta sqrt d), and the FP32 accumulation of 3d exact products adds
// gamma_3d (1 + 2^-11 + eta sqrt d)^2 (u = 2^-24: round-to-nearest accumulation, as fp16_eps).
double fp16x3_eps(uint32_t d) {
  const double u32 = std::ldexp(1.0, -24), u16 = std::ldexp(1.0, -11), uacc = u32;
  const double a = std::ldexp(1.0, -22) * (1.0 + u32) + u32;
  const double e = std::ldexp(1.0, -25) * std::sqrt((double)d);
  const double gam = 3.0 * d * uacc / (1.0 - 3.0 * d * uacc);
  return (a + e) * (2.0 + a + e) + (u16 + e) * (u16 + e) + gam * (1.0 + u16 + e) * (1.0 + u16 + e);
}

// Capacity of one candidate list.  A query's candidates are split over two lists (the two lanes of
// its lane pair, each seeing half of every sub-tile) and over the candidate splits; with K' = k +
// margin candidates per row (the reference's heap size), each list keeps max(k, ceil(K' / 2)): the
// two lists of a row together hold at least K', and all of its k nearest fit in either one (the
// default margin k gives lists of k).  The capacity only trades inserts against fallbacks; the
// certificate makes every choice exact.
uint32_t list_capacity(uint64_t n, uint64_t k, uint64_t margin) {
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(k, (k + margin + 1) / 2), n - 1);
}

int list_template(uint32_t kp) { return kp <= 16 ? 16 : kp <= 32 ? 32 : 64; }

Plan make_plan(const Device& D, uint64_t n, uint64_t d, uint64_t k, const costal_knn_params& p) {
  Plan P{};
  const uint64_t margin = p.margin < 0 ? k : (uint64_t)p.margin;
  P.kp = list_capacity(n, k, margin);
  // Tanimoto selection scans split FP16 rows whatever the precision asked for (the coarse form
  // never changes the output)
  P.tani = p.select == COSTAL_SELECT_TANIMOTO;
  P.prec = P.tani ? PREC_F16X3 : (int)p.precision;
  P.dp32 = round_up((uint32_t)std::max<uint64_t>(d, 1), 8);
  P.dp16 = round_up((uint32_t)std::max<uint64_t>(d, 1), 16);
  const uint32_t dmax = P.prec == PREC_F16 ? 128 : 64;
  P.fused = d <= dmax && P.kp <= FUSED_KP_MAX;
  P.J = P.prec == PREC_F16 ? P.dp16 / 16 : P.prec == PREC_F16X3 ? P.dp16 / 8 : P.dp32 / 8;
  P.KP = list_template(P.kp);
  P.NW = 4;
  P.QB = 32 * P.NW;
  P.TC = P.J <= 2 ? 512 : P.J <= 4 ? 256 : 128;  // TileShape<J>::TC
  P.n_pad = round_up((uint32_t)n, 512);  // whole tiles of every kind can be read
  const uint32_t ntile = (uint32_t)((n + P.TC - 1) / P.TC);
  const uint32_t nblk = (uint32_t)((n + P.QB - 1) / P.QB);
  uint32_t ns = p.col_splits;
  if (!ns) {
    const uint32_t target = 4u * (uint32_t)std::max(D.cus, 1);
    ns = (target + nblk - 1) / nblk;
    ns = std::min<uint32_t>(ns, 8);
  }
  ns = std::max<uint32_t>(1, std::min<uint32_t>(ns, std::max<uint32_t>(1, ntile / 2)));
  // the rescoring gathers 2 * splits * K' list entries per row into at most 2048 LDS slots
  ns = std::min<uint32_t>(ns, std::max<uint32_t>(1, RESCORE_CMAX / (2 * P.kp)));
  const uint32_t tiles_per = (ntile + ns - 1) / ns;
  P.cols_per_split = tiles_per * P.TC;
  P.nsplit = (ntile + tiles_per - 1) / tiles_per;
  if (P.nsplit < 1) P.nsplit = 1;
  // Pruned first pass: the splits take turns over each block's tiles in bound order (at most one
  // split per candidate tile of PRUNE_TILE rows).
  P.prune = false;
  if (P.fused && !p.no_prune) {
    P.order = costal::order_shape((uint32_t)n, P.dp32, P.QB, p.prune_clusters, P.tani, p.prune_tile);
    if (costal::order_supported(P.order)) {
      P.prune = true;
      P.nsplit = std::min<uint32_t>(P.nsplit, P.order.npt);
    }
  }
  P.wide = !P.fused && p.select == COSTAL_SELECT_COSINE && P.kp <= FUSED_KP_MAX;
  if (P.wide) {
    // rows of a multiple of 32 features; score blocks of up to 512 MB; the scan's candidate
    // sub-tiles split so that about four workgroups per CU run
    P.dp32 = round_up((uint32_t)std::max<uint64_t>(d, 1), WIDE_KC);
    P.n_pad = round_up((uint32_t)n, 256);  // whole gemm tiles (128 candidates) and scan blocks
    P.QB = WIDE_QB;
    P.TC = 32;
    const uint64_t row_bytes = (uint64_t)P.n_pad * 4;
    const uint64_t fit = std::max<uint64_t>(256, (512ull << 20) / row_bytes / 256 * 256);
    P.wide_qb = (uint32_t)std::min<uint64_t>(P.n_pad, fit);
    const uint32_t ncs = P.n_pad / 32;
    const uint32_t blocks = (std::min<uint32_t>(P.wide_qb, (uint32_t)n) + WIDE_QB - 1) / WIDE_QB;
    // every split adds 2 kp candidates per row to the f64 rescoring, which costs d multiply-adds
    // each, while a split's scan only reads stored scores: few splits
    uint32_t ws = p.col_splits ? p.col_splits : (4u * (uint32_t)std::max(D.cus, 1) + blocks - 1) / blocks;
    ws = std::min<uint32_t>(ws, p.col_splits ? 2048 / (2 * P.kp) : 2);
    ws = std::max<uint32_t>(1, std::min<uint32_t>(ws, ncs));
    const uint32_t per = (ncs + ws - 1) / ws;
    P.cols_per_split = per;  // candidate sub-tiles per split
    P.nsplit = (ncs + per - 1) / per;
  }
  return P;
}

int check_args(uint64_t n, uint64_t d, uint64_t k) {
  if (n > 0xffffffffull) return fail(COSTAL_E_TOO_MANY_ROWS, "n=%llu exceeds the u32 index range", (unsigned long long)n);
  if (k == 0 || k >= n)
    return fail(COSTAL_E_BAD_K, "invalid k=%llu: need 1 <= k <= n-1 (n=%llu)", (unsigned long long)k,
                (unsigned long long)n);
  if (d > 0xffffffffull / 8) return fail(COSTAL_E_UNSUPPORTED, "d=%llu is too large", (unsigned long long)d);
  return 0;
}

// Launch the exact row scan over `rows` (device list, count on device or host).
int launch_rowscan(Device& D, hipStream_t s, bool coarse, const float* x, uint32_t d, const float* xs32,
                   uint32_t dp32, const double* norms, uint32_t n, uint32_t k, uint32_t select,
                   const uint32_t* rows, const uint32_t* nrows_dev, uint32_t nrows_host, const double* lo,
                   double eps32, uint32_t* idx, double* cos, double* wts, uint32_t grid) {
  if (k > ROWSCAN_K_MAX)
    return fail(COSTAL_E_UNSUPPORTED, "k=%u above %u is not supported by the row-scan path", k, ROWSCAN_K_MAX);
  if (grid == 0) return 0;
  const bool small = k <= 512;
  if (coarse) {
    if (small)
      rowscan_kernel<1024, true><<<grid, 256, 0, s>>>(x, d, xs32, dp32, norms, n, k, select, rows, nrows_dev,
                                                      nrows_host, lo, eps32, idx, cos, wts);
    else
      rowscan_kernel<4096, true><<<grid, 256, 0, s>>>(x, d, xs32, dp32, norms, n, k, select, rows, nrows_dev,
                                                      nrows_host, lo, eps32, idx, cos, wts);
  } else {
    if (small)
      rowscan_kernel<1024, false><<<grid, 256, 0, s>>>(x, d, xs32, dp32, norms, n, k, select, rows, nrows_dev,
                                                       nrows_host, lo, eps32, idx, cos, wts);
    else
      rowscan_kernel<4096, false><<<grid, 256, 0, s>>>(x, d, xs32, dp32, norms, n, k, select, rows, nrows_dev,
                                                       nrows_host, lo, eps32, idx, cos, wts);
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

// A caller-fixed workspace (base, capacity) in place of the grow-only one.
struct FixedWs {
  void* base;
  size_t cap;
};

// The whole path on device pointers.  Serialized by the caller (D.mu held).  With `fixed`, every
// buffer comes from that region and nothing is allocated: a plan that does not fit is redone with
// one candidate range per row (and, for wide rows, smaller score blocks), then refused.
int knn_device_locked(Device& D, const float* x, uint64_t n64, uint64_t d64, const costal_knn_params& p,
                      uint32_t* idx, double* cos, double* norms_out, double* wts, costal_knn_stats* st,
                      hipStream_t s, bool brute, const FixedWs* fixed = nullptr) {
  if (st) {
    memset(st, 0, sizeof *st);
    st->nonfinite_row = st->nonfinite_sample = UINT64_MAX;
  }
  const uint32_t n = (uint32_t)n64, d = (uint32_t)d64, k = (uint32_t)p.k;
  if (p.precision > COSTAL_PREC_FP16X3) return fail(COSTAL_E_INVALID, "unknown precision %u", p.precision);
  if (p.select > COSTAL_SELECT_TANIMOTO) return fail(COSTAL_E_INVALID, "unknown select %u", p.select);
  if (p.prune_tile != 0 && p.prune_tile != 256 && p.prune_tile != 512)
    return fail(COSTAL_E_INVALID, "prune_tile=%u: need 0, 256 or 512", p.prune_tile);
  if (p.select == COSTAL_SELECT_TANIMOTO && !brute && (p.row_threshold || p.mlp_params))
    return fail(COSTAL_E_UNSUPPORTED, "admission thresholds are cosine similarities: Tanimoto selection takes none");
  Plan P = make_plan(D, n, d, k, p);
  if (brute) P.fused = P.wide = false;
  // Balanced first pass (tuned splits only): with at least one query block per resident workgroup
  // slot, the slots take equal shares of the (query block, tile) units instead of whole blocks,
  // so no slot idles while the last blocks finish; a block then has at most two pieces.  A second
  // piece starts with empty lists (more inserts) and the workgroups no longer stream the same tiles
  // together, so this pays where the scan itself is long against its inserts: the FP32 form, and
  // split FP16 with long lists (measured on 86,864 x 39 and 396,460 x 15; FP16 and split FP16 with
  // k = 10 lists were slower).
  uint32_t units_per_wg = 0, grid_bal = 0;
  if (P.fused && !P.prune && p.col_splits == 0 && (P.prec == PREC_F32 || (P.prec == PREC_F16X3 && P.KP >= 32))) {
    CoarseFn f0 = coarse_select(P.prec, P.J, P.KP, P.tani);
    int per_cu = 0;
    if (f0 && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(f0),
                                                           64 * P.NW, 0) == hipSuccess && per_cu > 0) {
      const uint64_t slots = (uint64_t)per_cu * (uint64_t)std::max(D.cus, 1);
      const uint64_t nblk = ((uint64_t)n + P.QB - 1) / P.QB, nt = ((uint64_t)n + P.TC - 1) / P.TC;
      const uint64_t units = nblk * nt;
      const uint64_t per = std::max<uint64_t>((units + slots - 1) / slots, nt);
      if (nblk >= slots && per <= 0xffffffffull) {
        units_per_wg = (uint32_t)per;
        grid_bal = (uint32_t)((units + per - 1) / per);
        P.nsplit = 2;
      }
    }
    (void)hipGetLastError();
  }
  const bool need16 = P.fused && (P.prec == PREC_F16 || P.prec == PREC_F16X3);
  const int split16 = P.prec == PREC_F16X3 ? 1 : 0;
  const size_t row16 = (size_t)P.dp16 * (split16 ? 2 : 1);
  // list capacity of the rescan of first-pass failures; the list buffers hold both passes' lists
  // (the rescan keeps at most 2 lists of kp2 per failed row per split, and at least one split fits)
  const uint32_t kp2 = (uint32_t)std::min<uint64_t>(
      std::max<uint64_t>(P.kp, std::min<uint64_t>((uint64_t)k + (k + 1) / 2, FUSED_KP_MAX)), n - 1);
  size_t nq_slots = 0, nq_entries = 0, score_floats = 0;
  const bool learned = P.fused && p.mlp_params && !p.row_threshold;
  // The rescoring gathers candidate rows 16 bytes at a time: from the input itself when its rows
  // are whole 16-byte words, else from a copy padded to dp32 floats that prep writes (left out,
  // and the rows read one float at a time, when a caller's fixed scratch cannot hold it).
  const bool x_v4 = d % 4 == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  bool xr_copy = (P.fused || P.wide) && !x_v4;
  auto workspace_bytes = [&]() -> size_t {
    // two candidate lists per query and split (one per lane of the query's lane pair)
    nq_slots = (P.fused || P.wide) ? (size_t)n * P.nsplit * 2 : 0;
    nq_entries = P.fused ? std::max<size_t>(nq_slots * P.kp, (size_t)n * 2 * kp2) : nq_slots * P.kp;
    score_floats = P.wide ? (size_t)P.wide_qb * P.n_pad : 0;
    Carve c(nullptr);
    c.take<float>((size_t)P.n_pad * P.dp32);
    if (need16) c.take<_Float16>((size_t)P.n_pad * row16);
    if (P.prune) {
      if (!need16 || P.prec == PREC_F16) c.take<float>((size_t)P.n_pad * P.dp32);
      c.take<uint8_t>(costal::order_bytes(P.order));
    }
    c.take<double>(n);
    c.take<uint32_t>(nq_entries);
    c.take<float>(nq_entries);
    c.take<uint32_t>(nq_slots);
    c.take<float>(nq_slots);
    c.take<uint32_t>(n);
    c.take<double>(n);
    c.take<uint32_t>(n);
    c.take<double>(n);
    if (P.fused) c.take<uint32_t>(n);
    c.take<uint32_t>(COUNTER_WORDS);
    if (learned) c.take<float>(n);
    if (P.fused) c.take<float>(n);
    c.take<float>(score_floats);
    if (P.fused && P.tani) {
      c.take<float>(P.n_pad);
      c.take<float2>(P.n_pad / (P.TC / 2));
    }
    if (xr_copy) c.take<float>((size_t)P.n_pad * P.dp32);
    return c.off + 256;
  };
  size_t bytes = workspace_bytes();
  if (fixed && bytes > fixed->cap && xr_copy) {
    xr_copy = false;  // rows read one float at a time
    bytes = workspace_bytes();
  }
  if (fixed && bytes > fixed->cap && P.prune) {
    P.prune = false;  // the bound matrix does not fit: scan every tile
    bytes = workspace_bytes();
  }
  if (fixed && bytes > fixed->cap && (P.fused || P.wide)) {
    costal_knn_params p1 = p;
    p1.col_splits = 1;
    const Plan P1 = make_plan(D, n, d, k, p1);
    P.nsplit = P1.nsplit;
    P.cols_per_split = P1.cols_per_split;
    units_per_wg = grid_bal = 0;
    bytes = workspace_bytes();
    while (P.wide && bytes > fixed->cap && P.wide_qb > 256) {
      P.wide_qb = std::max<uint32_t>(256, P.wide_qb / 2 / 256 * 256);
      bytes = workspace_bytes();
    }
  }
  int rc;
  if (fixed) {
    if (bytes > fixed->cap)
      return fail(COSTAL_E_UNSUPPORTED,
                  "n=%u, d=%u, k=%u needs %zu bytes of device scratch; the library's static scratch holds %zu",
                  n, d, k, bytes, fixed->cap);
  } else if ((rc = ws_reserve(D, bytes))) {
    return rc;
  }
  Carve c(fixed ? fixed->base : D.ws);
  float* xs32 = c.take<float>((size_t)P.n_pad * P.dp32);
  _Float16* xs16 = need16 ? c.take<_Float16>((size_t)P.n_pad * row16) : nullptr;
  // pruned pass: xs16 holds the working order; the f32 copy in working order when the scan reads it
  float* xs32w = nullptr;
  costal::OrderBufs ob{};
  if (P.prune) {
    if (!need16 || P.prec == PREC_F16) xs32w = c.take<float>((size_t)P.n_pad * P.dp32);
    ob = costal::order_carve(c.take<uint8_t>(costal::order_bytes(P.order)), P.order);
  }
  double* norms_ws = c.take<double>(n);
  uint32_t* cand_idx = c.take<uint32_t>(nq_entries);
  float* cand_sc = c.take<float>(nq_entries);
  uint32_t* cand_cnt = c.take<uint32_t>(nq_slots);
  float* cand_floor = c.take<float>(nq_slots);
  uint32_t* fb_rows = c.take<uint32_t>(n);
  double* fb_lo = c.take<double>(n);
  uint32_t* fb2_rows = c.take<uint32_t>(n);
  double* fb2_lo = c.take<double>(n);
  uint32_t* fb3_rows = P.fused ? c.take<uint32_t>(n) : nullptr;
  // [0] first bad row, [1] rows failing the first scan, [2] of them with < k admitted, [3] rows
  // failing the rescans, [4] short rows of them, [5] and [6] those with a k-th score and the
  // short ones (after a second-chance rescan), [7] rows failing the rescan of the short ones,
  // [12] short rows of that, [8..9] candidate-list inserts (u64), [10..11] tiles scanned (u64)
  uint32_t* counters = c.take<uint32_t>(COUNTER_WORDS);
  float* th_ws = learned ? c.take<float>(n) : nullptr;
  float* th_rescan = P.fused ? c.take<float>(n) : nullptr;
  float* scores = P.wide ? c.take<float>(score_floats) : nullptr;
  float* tnorm = (P.fused && P.tani) ? c.take<float>(P.n_pad) : nullptr;
  float2* trange = (P.fused && P.tani) ? c.take<float2>(P.n_pad / (P.TC / 2)) : nullptr;
  float* xr = xr_copy ? c.take<float>((size_t)P.n_pad * P.dp32) : nullptr;
  double* norms = norms_out ? norms_out : norms_ws;
  // the rows the rescoring reads, and their stride in floats
  const bool rows_v4 = x_v4 || xr_copy;
  const float* xrows = xr_copy ? xr : x;
  const uint32_t rstride = xr_copy ? P.dp32 : d;

  HIP_TRY(hipEventRecord(D.ev0, s));
  // The learned thresholds only read x: they run on the side stream from the start (behind
  // whatever the caller queued on s before the call), beside the prep and the locality order, and
  // join s before their first reader (the bound sort, or the scan).
  const bool learned_early = P.fused && learned;
  if (learned_early) {
    HIP_TRY(hipEventRecord(D.ev_fork, s));
    HIP_TRY(hipStreamWaitEvent(D.side, D.ev_fork, 0));
    // lowered by twice the list error bound, so that a row whose k-th similarity the model bounds
    // correctly also clears the certificate
    const double eps_l = P.prec == PREC_F16X3 ? fp16x3_eps(d) : fp32_eps(d);
    rc = costal_mlp_predict_device(x, n, d, p.mlp_params, p.mlp_bf16, p.mlp_offset, (float)(2.0 * eps_l), 0, th_ws,
                                   D.side);
    HIP_TRY(hipEventRecord(D.ev_join, D.side));
    if (rc) {
      (void)hipStreamSynchronize(D.side);
      return fail(rc, "threshold model: %s", costal_mlp_last_error());
    }
  }
  counters_init_kernel<<<1, 256, 0, s>>>(counters, COUNTER_WORDS, 1);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(D.evk[0], s));
  prep_kernel<<<(P.n_pad + PREP_ROWS - 1) / PREP_ROWS, 256, 0, s>>>(x, n, d, P.n_pad, P.dp32, P.dp16, split16,
                                                                         norms, xs32, P.prune ? nullptr : xs16,
                                                                         xr, counters);
  HIP_TRY(hipGetLastError());
  // Tanimoto: the norm extremes come back with the non-finite check (counters[14..17])
  if (P.fused && P.tani) {
    norm_extremes_kernel<<<std::min<uint32_t>((n + 255) / 256, 256), 256, 0, s>>>(
        norms, n, reinterpret_cast<unsigned long long*>(counters + 14));
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(D.evk[1], s));
  uint32_t head[18] = {};
  if ((rc = readback(D, head, counters, (P.fused && P.tani ? 18 : 1) * sizeof(uint32_t), s))) return rc;
  const uint32_t bad = head[0];
  if (bad != NONE) {
    if (learned_early) HIP_TRY(hipStreamSynchronize(D.side));  // nothing still reads x when we return
    if (st) st->nonfinite_row = bad;
    return fail(COSTAL_E_NON_FINITE, "row %u contains NaN or inf", bad);
  }
  // Tanimoto: norms scaled by a power of two that puts the largest in [1, 2); a spread of norms
  // beyond 2^100 (the smallest would lose bits in f32) takes the exact row scan instead
  double tani_scale = 1.0, norm_lo = 0.0, norm_hi = 0.0;
  if (P.fused && P.tani) {
    const uint64_t hib = (uint64_t)head[14] | ((uint64_t)head[15] << 32);
    const uint64_t lob = (uint64_t)head[16] | ((uint64_t)head[17] << 32);
    if (hib) {
      const uint64_t lbits = ~lob;
      memcpy(&norm_hi, &hib, sizeof norm_hi);
      memcpy(&norm_lo, &lbits, sizeof norm_lo);
      tani_scale = std::ldexp(1.0, -std::ilogb(norm_hi));
      if (norm_lo * tani_scale < std::ldexp(1.0, -100)) P.fused = P.prune = false;
    }
  }
  if (P.prune) {
    // the locality order, the working copies in it, and every block's tiles in bound order
    HIP_TRY(hipEventRecord(D.evk[6], s));
    HIP_TRY(costal::order_rows(s, xs32, P.order, ob, norms, norm_lo, norm_hi));
    HIP_TRY(costal::order_permute(s, xs32, P.order, ob, P.n_pad, P.dp16, split16, xs32w, xs16));
    HIP_TRY(costal::order_bounds(s, xs32, P.order, ob, norms));
    HIP_TRY(hipEventRecord(D.evk[7], s));
  }
  if (learned_early) HIP_TRY(hipStreamWaitEvent(s, D.ev_join, 0));  // the scan reads them
  if (P.fused && P.tani) {
    tani_norms_kernel<<<P.n_pad / (P.TC / 2), 64, 0, s>>>(norms, P.prune ? ob.perm : nullptr, n, P.TC / 2, tani_scale,
                                                         tnorm, trange);
    HIP_TRY(hipGetLastError());
  }

  const double eps32 = fp32_eps(d);
  // the row scan's f32 survivor test: cosines, or Tanimoto scores of them
  const double eps_row = P.tani ? tani_eps(eps32) : eps32;
  const uint32_t sel = P.fused && P.tani ? COSTAL_SELECT_TANIMOTO : COSTAL_SELECT_COSINE;
  unsigned long long* inserts = reinterpret_cast<unsigned long long*>(counters + 8);
  unsigned long long* err_max = reinterpret_cast<unsigned long long*>(counters + 64);  // ERR_SLOTS
  // f64 rescoring of the candidate lists of the queries `rows` (null: every row) with the
  // certificate; rows that fail it are appended to (fbr, fbl) and counted in *fbc
  auto rescore = [&](uint32_t kp, uint32_t nsplit, const uint32_t* rows, uint32_t nq, int all_admitted,
                     double eps_list, uint32_t* fbr, double* fbl, uint32_t* fbc, uint32_t* shc) -> int {
    const uint32_t nlists = 2 * nsplit;
    const uint32_t slots = nlists * kp;
    // rows_v4: the rows are read 16 bytes at a time (xrows, rstride as dot64r<true> reads them)
    auto launch = [&](auto v4) {
      constexpr bool V4 = decltype(v4)::value;
#define RESCORE(CM)                                                                                       \
  rescore_kernel<CM, V4><<<nq, rescore_threads(CM), 0, s>>>(                                                  \
      xrows, P.prune ? ob.perm : nullptr, rstride, norms, n, k, (int)kp, nlists, rows, nq, cand_idx, cand_sc, \
      cand_cnt, cand_floor, all_admitted, eps_list, eps_row, sel, idx, cos, wts, fbr, fbl, fbc, shc, err_max)
      if (slots <= 32) {
        rescore_pair_kernel<V4><<<(nq + 1) / 2, 64, 0, s>>>(xrows, P.prune ? ob.perm : nullptr, rstride, norms, k,
                                                            (int)kp, nlists, rows, nq, cand_idx, cand_sc, cand_cnt,
                                                            cand_floor, all_admitted, eps_list, eps_row, sel, idx, cos,
                                                            wts, fbr, fbl, fbc, shc, err_max);
      } else if (slots <= 64) RESCORE(64);
      else if (slots <= 128) RESCORE(128);
      else if (slots <= 256) RESCORE(256);
      else if (slots <= 512) RESCORE(512);
      else if (slots <= 1024) RESCORE(1024);
      else if (slots <= 2048) RESCORE(2048);
      else RESCORE(4096);
#undef RESCORE
    };
    if (rows_v4) launch(std::true_type{});
    else launch(std::false_type{});
    HIP_TRY(hipGetLastError());
    return 0;
  };
  bool third = false;  // rows left short by a second-chance rescan were rescanned once more
  if (P.fused) {
    CoarseFn fn = coarse_select(P.prec, P.J, P.KP, P.tani);
    if (!fn) return fail(COSTAL_E_UNSUPPORTED, "no coarse kernel for d=%u", d);
    const float eps_pre = P.prec == PREC_F16 ? (float)(fp16_eps(d) + eps32 + std::ldexp(1.0, -22)) : 0.0f;
    const uint8_t* xsq = need16 ? (const uint8_t*)xs16 : (const uint8_t*)(xs32w ? xs32w : xs32);
    const float* xs32q = xs32w ? xs32w : xs32;  // the f32 rows the FP16 form re-scores, in scan order
    const double eps_list = P.tani ? tani_eps(fp16x3_eps(d)) : P.prec == PREC_F16X3 ? fp16x3_eps(d) : eps32;
    const float* row_th = p.row_threshold;
    unsigned long long* scanned = reinterpret_cast<unsigned long long*>(counters + 10);
    // the first pass prunes; the rescans of failed rows read the same working copies, unpruned
    ScanArgs pr_first{}, pr_rescan{};
    if (P.prune) {
      pr_rescan.perm = ob.perm;
      pr_rescan.iperm = ob.iperm;
      pr_first = pr_rescan;
      pr_first.seq = ob.seq;
      pr_first.sb = ob.sb;
      pr_first.npt = P.order.npt;
      pr_first.pt = P.order.pt;
      pr_first.scanned = scanned;
    }
    pr_first.tnorm = pr_rescan.tnorm = tnorm;
    pr_first.trange = pr_rescan.trange = trange;
    if (learned) row_th = th_ws;  // predicted behind the non-finite check
    const uint32_t ntile = (n + P.TC - 1) / P.TC;
    // one scan pass over the queries `rows` (null: every row), then rescoring; rows that fail the
    // certificate are appended to (fbr, fbl) and counted in *fbc
    auto scan_pass = [&](CoarseFn f, uint32_t kp, const uint32_t* rows, uint32_t nq, const float* th,
                         uint32_t nsplit, uint32_t cols_per, uint32_t units, uint32_t* fbr, double* fbl,
                         uint32_t* fbc, uint32_t* shc, const ScanArgs& pr) -> int {
      const dim3 grid = units ? dim3(grid_bal, 1) : dim3((nq + P.QB - 1) / P.QB, nsplit);
      f<<<grid, 64 * P.NW, 0, s>>>(xsq, xs32q, P.dp32, n, (int)kp, rows, nq, th, cols_per, units, eps_pre,
                                   cand_idx, cand_sc, cand_cnt, cand_floor, inserts, pr);
      HIP_TRY(hipGetLastError());
      if (!rows) HIP_TRY(hipEventRecord(D.evk[3], s));
      const int all_admitted = (nsplit == 1 && kp >= n - 1 && !th) ? 1 : 0;
      return rescore(kp, nsplit, rows, nq, all_admitted, eps_list, fbr, fbl, fbc, shc);
    };
    HIP_TRY(hipEventRecord(D.evk[2], s));
    rc = scan_pass(fn, P.kp, nullptr, n, row_th, P.nsplit, P.cols_per_split, units_per_wg, fb_rows, fb_lo,
                   counters + 1, counters + 2, pr_first);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(D.evk[4], s));
    const uint32_t grid_fb = std::min<uint32_t>(n, 4u * (uint32_t)std::max(D.cus, 1));
    // Rows that failed the first pass (admission thresholds above their k-th similarity, or lists
    // too short for their ties) are rescanned with longer lists as one batched scan over the row
    // list (wide candidate splits keep the few rows' scan parallel); only rows that fail again
    // (exact ties) go to the full-row kernel.
    uint32_t nfb_short[2] = {0, 0};  // rows that failed, and those of them with fewer than k candidates
    if ((rc = readback(D, nfb_short, counters + 1, sizeof nfb_short, s))) return rc;
    const uint32_t nfb = nfb_short[0], nshort = nfb_short[1];
    if (nfb > 0) {
      CoarseFn fn2 = coarse_select(P.prec, P.J, list_template(kp2), P.tani);
      if (!fn2) return fail(COSTAL_E_UNSUPPORTED, "no coarse kernel for d=%u", d);
      // One rescan of every failed row, with lists of kp2 for near-ties.  A row that failed only
      // the certificate has its k-th exact score s_k among the first pass's candidates (fb_lo =
      // s_k - eps32), and its true k-th is at least that: it admits only scores above fb_lo -
      // eps_list, which every true neighbor's list score exceeds.  A row with fewer than k
      // candidates (its admission threshold was too high) rescans with a second, lower learned
      // threshold when the model has one (so that the rescan still prunes), else with none.  The
      // thresholds are indexed by row, so both kinds share the launch.
      rescan_thresholds_kernel<<<(nfb + 255) / 256, 256, 0, s>>>(fb_rows, fb_lo, nfb, eps_list, th_rescan);
      HIP_TRY(hipGetLastError());
      const bool chance = nshort > 0 && learned && p.mlp_offset2 != 0.0f && p.mlp_offset2 > p.mlp_offset;
      if (chance) {
        const float margin = (float)(2.0 * eps_list);
        const float factor = (float)std::exp((double)p.mlp_offset2 - (double)p.mlp_offset);
        second_chance_kernel<<<(nfb + 255) / 256, 256, 0, s>>>(fb_rows, fb_lo, nfb, th_ws, margin, factor, th_rescan);
        HIP_TRY(hipGetLastError());
      }
      // max_splits, and list entries per row for the rescoring (2 * splits * kpx at most max_entries)
      auto rescan = [&](CoarseFn f, uint32_t kpx, uint32_t* rows, uint32_t nr, const float* th, uint32_t* outr,
                        double* outl, uint32_t* outc, uint32_t* outs, uint32_t max_splits,
                        uint32_t max_entries) -> int {
        // about two workgroups per CU: the rows are few, so splits are what fills the GPU, but
        // every split fills its lists from scratch (10^6 x 40, k = 30, learned: one per CU 8.3 ms
        // for the fallback, two 7.1 ms; 396,460 x 15: 1.00 -> 0.95 ms)
        const bool pruned = P.prune && nr <= costal::RESCAN_PRUNE_ROWS;
        const uint32_t blocks = (nr + P.QB - 1) / P.QB;
        uint32_t ns = (2u * (uint32_t)std::max(D.cus, 1) + blocks - 1) / blocks;
        ns = std::min<uint32_t>(ns, max_splits);
        ns = std::min<uint32_t>(ns, std::max<uint32_t>(1, ntile / 2));
        ns = std::min<uint32_t>(ns, std::max<uint32_t>(1, max_entries / (2 * kpx)));
        // the lists fit in the list buffers (nq_entries entries, nq_slots lists; one split always
        // fits since nr <= n and kpx <= kp2)
        ns = std::min<uint64_t>(ns, nq_entries / ((uint64_t)nr * 2 * kpx));
        ns = std::min<uint64_t>(ns, nq_slots / ((uint64_t)nr * 2));
        ns = std::max<uint32_t>(ns, 1);
        const uint32_t tiles_per = (ntile + ns - 1) / ns;
        ns = (ntile + tiles_per - 1) / tiles_per;
        // In the locality order the rows are sorted into working order, so that a block holds
        // similar rows, and get cones and bounds of their own: the rescan prunes like the first pass.
        ScanArgs pr = pr_rescan;
        if (pruned) {
          HIP_TRY(costal::order_rescan(s, xs32, P.order, ob, rows, nr, norms));
          pr.seq = ob.rseq;
          pr.sb = ob.rsb;
          pr.npt = P.order.npt;
          pr.pt = P.order.pt;
          ns = std::min<uint32_t>(ns, P.order.npt);
        }
        return scan_pass(f, kpx, rows, nr, th, ns, tiles_per * P.TC, 0, outr, outl, outc, outs, pr);
      };
      // Lists of kp2 are for rows with near-ties (396,460 x 15: with lists of 10 instead of 15, three
      // rows went on to the row scan, 1.3 ms more).  They cost nothing where they fit the first
      // pass's list capacity; otherwise, where short rows are most of the failures (learned
      // thresholds), the first pass's lists serve (10^6 x 40, k = 30: 5,038 of 5,063 rows short,
      // 10.6 ms with lists of 45, whose kernel runs one wave per SIMD).
      const bool long_lists = list_template(kp2) == P.KP || 2 * nshort < nfb;
      if ((rc = rescan(long_lists ? fn2 : fn, long_lists ? kp2 : P.kp, fb_rows, nfb, th_rescan, fb2_rows, fb2_lo,
                       counters + 3, counters + 4, 32, RESCORE_CMAX)))
        return rc;
      if (chance) {
        // Rows the second chance left short are few, and the latency of their scan is what
        // counts: a threshold-free rescan with up to 96 splits (up to 4096 list entries per row:
        // the rescoring takes the top k of many entries quickly), not the row scan (which would
        // stream every candidate of such a row through one workgroup).
        uint32_t c34[2] = {0, 0};
        if ((rc = readback(D, c34, counters + 3, sizeof c34, s))) return rc;
        if (c34[1] > 0) {
          split_failures_kernel<<<(c34[0] + 255) / 256, 256, 0, s>>>(fb2_rows, fb2_lo, c34[0], fb_rows, fb_lo,
                                                                    counters + 5, fb3_rows, counters + 6);
          HIP_TRY(hipGetLastError());
          if ((rc = rescan(fn, P.kp, fb3_rows, c34[1], nullptr, fb2_rows, fb2_lo, counters + 7, counters + 12,
                           96, RESCORE_CMAX_FEW)))
            return rc;
          third = true;
        }
      }
    }
    if (third) {
      rc = launch_rowscan(D, s, true, x, d, xs32, P.dp32, norms, n, k, sel, fb_rows, counters + 5, 0, fb_lo, eps_row,
                          idx, cos, wts, grid_fb);
      if (!rc)
        rc = launch_rowscan(D, s, true, x, d, xs32, P.dp32, norms, n, k, sel, fb2_rows, counters + 7, 0, fb2_lo,
                            eps_row, idx, cos, wts, grid_fb);
    } else {
      rc = launch_rowscan(D, s, true, x, d, xs32, P.dp32, norms, n, k, sel, fb2_rows, counters + 3, 0, fb2_lo, eps_row,
                          idx, cos, wts, grid_fb);
    }
    if (rc) return rc;
    HIP_TRY(hipEventRecord(D.evk[5], s));
  } else if (P.wide) {
    const uint32_t ncs = P.n_pad / 32;
    HIP_TRY(hipEventRecord(D.evk[2], s));
    for (uint32_t q0 = 0; q0 < n; q0 += P.wide_qb) {
      const uint32_t nqb = std::min<uint32_t>(P.wide_qb, n - q0);
      const dim3 g1(round_up(nqb, WIDE_QT) / WIDE_QT, P.n_pad / WIDE_CT);
      wide_gemm_kernel<<<g1, 256, 0, s>>>(xs32, P.dp32, q0, ncs, scores);
      HIP_TRY(hipGetLastError());
      const dim3 g2((nqb + WIDE_QB - 1) / WIDE_QB, P.nsplit);
#define WIDE_SCAN(KPR)                                                                                \
  wide_scan_kernel<KPR><<<g2, 256, 0, s>>>(scores, q0, nqb, ncs, n, (int)P.kp, P.nsplit, P.cols_per_split, \
                                           cand_idx, cand_sc, cand_cnt, cand_floor, inserts)
      if (P.KP == 16) WIDE_SCAN(16);
      else if (P.KP == 32) WIDE_SCAN(32);
      else WIDE_SCAN(64);
#undef WIDE_SCAN
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(D.evk[3], s));
    const int all_admitted = (P.nsplit == 1 && P.kp >= n - 1) ? 1 : 0;
    rc = rescore(P.kp, P.nsplit, nullptr, n, all_admitted, eps32, fb_rows, fb_lo, counters + 1, counters + 2);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(D.evk[4], s));
    rc = launch_rowscan(D, s, true, x, d, xs32, P.dp32, norms, n, k, COSTAL_SELECT_COSINE, fb_rows, counters + 1,
                        0, fb_lo, eps32, idx, cos, wts, std::min<uint32_t>(n, 4u * (uint32_t)std::max(D.cus, 1)));
    if (rc) return rc;
    HIP_TRY(hipEventRecord(D.evk[5], s));
  } else {
    const uint32_t grid = std::min<uint32_t>(n, 8u * (uint32_t)std::max(D.cus, 1));
    HIP_TRY(hipEventRecord(D.evk[2], s));
    HIP_TRY(hipEventRecord(D.evk[3], s));
    HIP_TRY(hipEventRecord(D.evk[4], s));
    rc = launch_rowscan(D, s, false, x, d, xs32, P.dp32, norms, n, k, p.select, nullptr, nullptr, n, nullptr,
                        eps32, idx, cos, wts, grid);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(D.evk[5], s));
  }
  HIP_TRY(hipEventRecord(D.ev1, s));
  uint32_t cnt[COUNTER_WORDS] = {};
  if ((rc = readback(D, cnt, counters, sizeof cnt, s))) return rc;
  if (st) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, D.ev0, D.ev1) != hipSuccess) ms = 0.0f;
    st->fallback_rows = (P.fused || P.wide) ? cnt[1] : 0;
    st->short_rows = cnt[2];
    st->row_block = (P.fused || P.wide) ? P.QB : 1;
    st->col_block = (P.fused || P.wide) ? P.TC : 256;
    if (P.prune) {
      const uint64_t total = (uint64_t)P.order.nblk * P.order.npt;
      const uint64_t done = (uint64_t)cnt[10] | ((uint64_t)cnt[11] << 32);
      st->blocks_total = total;
      st->blocks_skipped = total > done ? total - done : 0;
    } else {
      st->blocks_skipped = 0;
      st->blocks_total = (P.fused || P.wide) ? (uint64_t)((n + P.QB - 1) / P.QB) * ((n + P.TC - 1) / P.TC) : n;
    }
    st->candidates = P.kp;
    st->col_splits = (P.fused || P.wide) ? P.nsplit : 1;
    st->path = P.fused ? 0 : P.wide ? 2 : 1;
    st->inserts = (uint64_t)cnt[8] | ((uint64_t)cnt[9] << 32);
    st->rowscan_rows = P.fused ? (third ? cnt[5] + cnt[7] : cnt[3]) : P.wide ? cnt[1] : n;
    st->device_ms = ms;
    auto span = [&](int a, int b) {
      float v = 0.0f;
      return hipEventElapsedTime(&v, D.evk[a], D.evk[b]) == hipSuccess ? (double)v : 0.0;
    };
    st->prep_ms = span(0, 1);
    st->coarse_ms = span(2, 3);
    st->rescore_ms = span(3, 4);
    st->fallback_ms = span(4, 5);
    st->order_ms = P.prune ? span(6, 7) : 0.0;
    st->coarse_err_max = slots_max(cnt + 64);
    st->coarse_eps = P.fused ? (P.tani ? tani_eps(fp16x3_eps(d)) : P.prec == PREC_F16X3 ? fp16x3_eps(d) : fp32_eps(d))
                     : P.wide ? fp32_eps(d) : 0.0;
  }
  return 0;
}

hipStream_t pick_stream(Device& D, void* stream) { return stream ? (hipStream_t)stream : D.stream; }

// the shortest decimal that reads back as x (NaN, inf and -inf spelled as Rust prints them)
std::string shortest(double x) {
  // NaN and inf by their bits (the device pass of this file is built without NaN semantics)
  uint64_t bits;
  memcpy(&bits, &x, sizeof bits);
  if ((bits & 0x7ff0000000000000ull) == 0x7ff0000000000000ull) {
    if (bits & 0x000fffffffffffffull) return "NaN";
    return (bits >> 63) ? "-inf" : "inf";
  }
  char buf[64];
  for (int p = 1; p <= 17; ++p) {
    snprintf(buf, sizeof buf, "%.*g", p, x);
    if (strtod(buf, nullptr) == x) break;
  }
  return buf;
}

// ------------------------------------------------------------------------------------------------
// Batched samples (host side)

using costal::CoarseBatchFn;

CoarseBatchFn coarse_batch_select(int prec, uint32_t J, int KP) {
  using namespace costal;
  if (prec == PREC_F32)
    return KP == 16 ? scan_batch_fn_p0_k16(J) : KP == 32 ? scan_batch_fn_p0_k32(J) : scan_batch_fn_p0_k64(J);
  if (prec == PREC_F16)
    return KP == 16 ? scan_batch_fn_p1_k16(J) : KP == 32 ? scan_batch_fn_p1_k32(J) : scan_batch_fn_p1_k64(J);
  return KP == 16 ? scan_batch_fn_p2_k16(J) : KP == 32 ? scan_batch_fn_p2_k32(J) : scan_batch_fn_p2_k64(J);
}

struct BatchGraph {
  std::vector<uint8_t> key;  // samples + params the graph was captured for
  void* ws = nullptr;        // the batch path's own workspace (the graph points into it)
  size_t ws_cap = 0;
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr;
  bool use_graph = true;
  bool ready = false;
  std::function<hipError_t(hipStream_t)> launch;  // the batch's launch sequence
  // host copies of what the stats need
  uint32_t QB = 0, TC = 0, max_split = 0, max_kp = 0;
  double eps_list = 0.0;
  uint64_t blocks_total = 0;
  size_t off_counters = 0;
};
BatchGraph g_batches[MAX_DEVICES];

int knn_batched_locked(Device& D, const costal_sample* smp, uint32_t S, const costal_knn_params& p,
                       costal_knn_stats* st, hipStream_t s) {
  BatchGraph& g_batch = g_batches[D.id];
  const uint64_t k = p.k;
  if (p.select != COSTAL_SELECT_COSINE) return fail(COSTAL_E_UNSUPPORTED, "batched samples use cosine selection");
  if (p.precision > COSTAL_PREC_FP16X3) return fail(COSTAL_E_INVALID, "unknown precision %u", p.precision);
  if (p.row_threshold || p.mlp_params) return fail(COSTAL_E_UNSUPPORTED, "batched samples take no admission thresholds");
  uint32_t dmax = 1;
  uint64_t total_rows = 0;
  for (uint32_t i = 0; i < S; ++i) {
    const costal_sample& a = smp[i];
    if (!a.x || !a.idx || !a.cos || !a.norms) return fail(COSTAL_E_INVALID, "sample %u: null pointer", i);
    if (a.n > 0xffffffffull) return fail(COSTAL_E_TOO_MANY_ROWS, "sample %u: n=%llu exceeds the u32 index range", i,
                                         (unsigned long long)a.n);
    if (k == 0 || k >= a.n)
      return fail(COSTAL_E_BAD_K, "sample %u: invalid k=%llu: need 1 <= k <= n-1 (n=%llu)", i, (unsigned long long)k,
                  (unsigned long long)a.n);
    if (a.d == 0 || a.d > 64) return fail(COSTAL_E_UNSUPPORTED, "sample %u: batched samples need 1 <= d <= 64", i);
    dmax = std::max<uint32_t>(dmax, (uint32_t)a.d);
    total_rows += a.n;
  }
  if (total_rows > 0xffffffffull) return fail(COSTAL_E_TOO_MANY_ROWS, "batch exceeds the u32 row range");
  // reuse the captured graph when the batch layout is unchanged
  std::vector<uint8_t> key(sizeof(costal_sample) * S + sizeof(costal_knn_params) + sizeof(void*));
  memcpy(key.data(), smp, sizeof(costal_sample) * S);
  memcpy(key.data() + sizeof(costal_sample) * S, &p, sizeof p);
  memcpy(key.data() + sizeof(costal_sample) * S + sizeof p, &s, sizeof(void*));

  const int prec = (int)p.precision;
  const uint32_t dp32 = round_up(dmax, 8), dp16 = round_up(dmax, 16);
  const uint32_t J = prec == PREC_F16 ? dp16 / 16 : prec == PREC_F16X3 ? dp16 / 8 : dp32 / 8;
  const uint32_t TC = J <= 2 ? 512 : J <= 4 ? 256 : 128;  // TileShape<J>::TC
  const uint32_t QB = 128;
  const bool need16 = prec != PREC_F32;
  const int split16 = prec == PREC_F16X3 ? 1 : 0;
  const size_t row16 = (size_t)dp16 * (split16 ? 2 : 1);

  if (g_batch.ready && g_batch.key == key) {
    // replay
  } else {
    if (g_batch.exec) {
      (void)hipGraphExecDestroy(g_batch.exec);
      g_batch.exec = nullptr;
    }
    g_batch.ready = false;
    std::vector<BatchSample> bs(S);
    std::vector<uint32_t> pad_first(S), row_first(S), items;
    uint32_t pad = 0, base = 0, max_split = 1, max_kp = 1;
    uint64_t slots = 0, entries = 0, blocks_total = 0;
    // Untuned samples are cut into work items of about 8192 candidates (items of similar cost)
    // only when the batch is small; a batch of at least 2048 query blocks fills the GPU with whole
    // samples, and every split would fill its candidate lists again.
    uint64_t batch_blocks = 0;
    for (uint32_t i = 0; i < S; ++i) batch_blocks += (smp[i].n + QB - 1) / QB;
    const uint32_t target_cols = batch_blocks >= 2048 ? 0xffffffffu : 8192;
    for (uint32_t i = 0; i < S; ++i) {
      const costal_sample& a = smp[i];
      const uint32_t n = (uint32_t)a.n;
      const int64_t mg = a.margin >= 0 ? a.margin : (p.margin >= 0 ? p.margin : (int64_t)k);
      const uint32_t kp = list_capacity(n, k, (uint64_t)mg);
      if (kp > FUSED_KP_MAX) return fail(COSTAL_E_UNSUPPORTED, "sample %u: k + ceil(margin/2) above %u", i, FUSED_KP_MAX);
      const uint32_t ntile = (n + TC - 1) / TC;
      uint32_t ns = a.col_splits ? a.col_splits : (uint32_t)(((uint64_t)n + target_cols - 1) / target_cols);
      ns = std::max<uint32_t>(1, std::min<uint32_t>(ns, ntile));
      ns = std::min<uint32_t>(ns, std::max<uint32_t>(1, 2048 / (2 * kp)));
      const uint32_t tiles_per = (ntile + ns - 1) / ns;
      ns = (ntile + tiles_per - 1) / tiles_per;
      BatchSample& b = bs[i];
      b.x = a.x;
      b.n = n;
      b.d = (uint32_t)a.d;
      b.row_pad = pad;
      b.row_base = base;
      b.kp = (int)kp;
      b.nsplit = ns;
      b.cols_per_split = tiles_per * TC;
      b.slot_base = slots;
      b.entry_base = entries;
      b.idx = a.idx;
      b.cos = a.cos;
      b.norms = a.norms;
      b.wts = a.weights;
      pad_first[i] = pad;
      row_first[i] = base;
      const uint32_t nqb = (n + QB - 1) / QB;
      for (uint32_t qb = 0; qb < nqb; ++qb)
        for (uint32_t sp = 0; sp < ns; ++sp) {
          items.push_back(i);
          items.push_back(qb);
          items.push_back(sp);
        }
      blocks_total += (uint64_t)nqb * ntile;
      pad += round_up(n, 512);
      base += n;
      slots += (uint64_t)n * ns * 2;
      entries += (uint64_t)n * ns * 2 * kp;
      max_split = std::max(max_split, ns);
      max_kp = std::max(max_kp, kp);
    }
    // largest work first (an item's cost is its candidate count)
    const uint32_t nitems = (uint32_t)(items.size() / 3);
    std::vector<uint32_t> order(nitems);
    for (uint32_t e = 0; e < nitems; ++e) order[e] = e;
    auto cost = [&](uint32_t e) {
      const BatchSample& b = bs[items[3 * e]];
      const uint32_t c0 = items[3 * e + 2] * b.cols_per_split;
      return std::min<uint32_t>(b.cols_per_split, b.n > c0 ? b.n - c0 : 0);
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });
    std::vector<uint32_t> sorted(items.size());
    for (uint32_t e = 0; e < nitems; ++e)
      for (int c = 0; c < 3; ++c) sorted[3 * e + c] = items[3 * order[e] + c];

    const int KP = list_template(max_kp);
    CoarseBatchFn fn = coarse_batch_select(prec, J, KP);
    if (!fn) return fail(COSTAL_E_UNSUPPORTED, "no batched scan kernel for d=%u", dmax);

    size_t bytes;
    {
      Carve c(nullptr);
      c.take<float>((size_t)pad * dp32);
      if (need16) c.take<_Float16>((size_t)pad * row16);
      c.take<uint32_t>(entries);
      c.take<float>(entries);
      c.take<uint32_t>(slots);
      c.take<float>(slots);
      c.take<uint32_t>(total_rows);
      c.take<double>(total_rows);
      c.take<uint32_t>(COUNTER_WORDS);
      c.take<BatchSample>(S);
      c.take<uint32_t>(S);
      c.take<uint32_t>(S);
      c.take<uint32_t>(items.size());
      bytes = c.off + 256;
    }
    if (bytes > g_batch.ws_cap) {
      if (g_batch.ws) {
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipFree(g_batch.ws));
        g_batch.ws = nullptr;
        g_batch.ws_cap = 0;
      }
      const size_t cap = bytes + bytes / 8 + (1u << 20);
      HIP_TRY(hipMalloc(&g_batch.ws, cap));
      g_batch.ws_cap = cap;
    }
    Carve c(g_batch.ws);
    float* xs32 = c.take<float>((size_t)pad * dp32);
    _Float16* xs16 = need16 ? c.take<_Float16>((size_t)pad * row16) : nullptr;
    uint32_t* cand_idx = c.take<uint32_t>(entries);
    float* cand_sc = c.take<float>(entries);
    uint32_t* cand_cnt = c.take<uint32_t>(slots);
    float* cand_floor = c.take<float>(slots);
    uint32_t* fb_rows = c.take<uint32_t>(total_rows);
    double* fb_lo = c.take<double>(total_rows);
    uint32_t* counters = c.take<uint32_t>(COUNTER_WORDS);
    const size_t off_counters = (size_t)(reinterpret_cast<uint8_t*>(counters) - static_cast<uint8_t*>(g_batch.ws));
    BatchSample* d_samples = c.take<BatchSample>(S);
    uint32_t* d_pad_first = c.take<uint32_t>(S);
    uint32_t* d_row_first = c.take<uint32_t>(S);
    uint32_t* d_items = c.take<uint32_t>(items.size());
    HIP_TRY(hipMemcpyAsync(d_samples, bs.data(), sizeof(BatchSample) * S, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_pad_first, pad_first.data(), sizeof(uint32_t) * S, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_row_first, row_first.data(), sizeof(uint32_t) * S, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_items, sorted.data(), sizeof(uint32_t) * sorted.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));

    const double eps32 = fp32_eps(dmax);
    const double eps_list = prec == PREC_F16X3 ? fp16x3_eps(dmax) : eps32;
    const float eps_pre = prec == PREC_F16 ? (float)(fp16_eps(dmax) + eps32 + std::ldexp(1.0, -22)) : 0.0f;
    const uint8_t* xsq = need16 ? (const uint8_t*)xs16 : (const uint8_t*)xs32;
    const uint32_t max_slots = 2 * max_split * max_kp;
    const uint32_t grid_scan = 2u * (uint32_t)std::max(D.cus, 1);
    const uint32_t grid_fb = 4u * (uint32_t)std::max(D.cus, 1);
    // counters: [0..1] first bad (sample, row), [2] rows failing the certificate, [3] of them short,
    // [4..5] inserts, [6] work-queue head, [64..) ERR_SLOTS maxima of the measured list-score error
    const unsigned long long* bad = reinterpret_cast<const unsigned long long*>(counters);
    unsigned long long* err_max = reinterpret_cast<unsigned long long*>(counters + 64);

    // the whole batch as one sequence of launches (no host round trip, no memset nodes)
    g_batch.launch = [=](hipStream_t st) {
      batch_init_kernel<<<1, 256, 0, st>>>(counters, COUNTER_WORDS);
      prep_batch_kernel<<<(pad + 255) / 256, 256, 0, st>>>(d_samples, d_pad_first, S, pad, dp32, dp16, split16, xs32,
                                                           xs16, reinterpret_cast<unsigned long long*>(counters));
      fn<<<grid_scan, 256, 0, st>>>(d_samples, d_items, nitems, counters + 6, xsq, xs32, dp32, eps_pre, cand_idx,
                                    cand_sc, cand_cnt, cand_floor, reinterpret_cast<unsigned long long*>(counters + 4),
                                    bad);
#define RESCORE_B(CM)                                                                                             \
  rescore_batch_kernel<CM><<<(uint32_t)total_rows, 64, 0, st>>>(d_samples, d_row_first, S, (uint32_t)total_rows, \
                                                                (uint32_t)k, eps_list, eps32, cand_idx, cand_sc,  \
                                                                cand_cnt, cand_floor, fb_rows, fb_lo, counters + 2, \
                                                                counters + 3, err_max, bad)
      if (max_slots <= 32)
        rescore_batch_pair_kernel<<<(uint32_t)((total_rows + 1) / 2), 64, 0, st>>>(
            d_samples, d_row_first, S, (uint32_t)total_rows, (uint32_t)k, eps_list, eps32, cand_idx, cand_sc,
            cand_cnt, cand_floor, fb_rows, fb_lo, counters + 2, counters + 3, err_max, bad);
      else if (max_slots <= 64) RESCORE_B(64);
      else if (max_slots <= 128) RESCORE_B(128);
      else if (max_slots <= 256) RESCORE_B(256);
      else if (max_slots <= 512) RESCORE_B(512);
      else if (max_slots <= 1024) RESCORE_B(1024);
      else RESCORE_B(2048);
#undef RESCORE_B
      if (k <= 512)
        rowscan_batch_kernel<1024><<<grid_fb, 256, 0, st>>>(d_samples, d_row_first, S, xs32, dp32, (uint32_t)k,
                                                            fb_rows, counters + 2, fb_lo, eps32, bad);
      else
        rowscan_batch_kernel<4096><<<grid_fb, 256, 0, st>>>(d_samples, d_row_first, S, xs32, dp32, (uint32_t)k,
                                                            fb_rows, counters + 2, fb_lo, eps32, bad);
      return hipGetLastError();
    };
    g_batch.use_graph = getenv("COSTAL_NO_GRAPH") == nullptr;
    if (g_batch.use_graph) {
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipError_t le = g_batch.launch(s);
      hipError_t ce = hipStreamEndCapture(s, &graph);
      if (le != hipSuccess) return fail(COSTAL_E_DEVICE, "batched launch failed: %s", hipGetErrorString(le));
      if (ce != hipSuccess) return fail(COSTAL_E_DEVICE, "graph capture failed: %s", hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&g_batch.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        g_batch.exec = nullptr;
        return fail(COSTAL_E_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ie));
      }
    }
    g_batch.key = key;
    g_batch.ready = true;
    g_batch.stream = s;
    g_batch.QB = QB;
    g_batch.TC = TC;
    g_batch.max_split = max_split;
    g_batch.max_kp = max_kp;
    g_batch.eps_list = eps_list;
    g_batch.blocks_total = blocks_total;
    g_batch.off_counters = off_counters;
  }

  uint32_t* counters = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(g_batch.ws) + g_batch.off_counters);
  HIP_TRY(hipEventRecord(D.ev0, s));
  if (g_batch.use_graph) {
    HIP_TRY(hipGraphLaunch(g_batch.exec, s));
  } else {
    HIP_TRY(g_batch.launch(s));
  }
  HIP_TRY(hipEventRecord(D.ev1, s));
  uint32_t cnt[COUNTER_WORDS] = {};
  if (int rb = readback(D, cnt, counters, sizeof cnt, s)) return rb;
  if (cnt[0] != NONE || cnt[1] != NONE) {
    if (st) {
      memset(st, 0, sizeof *st);
      st->nonfinite_row = cnt[0];
      st->nonfinite_sample = cnt[1];
    }
    return fail(COSTAL_E_NON_FINITE, "sample %u: row %u contains NaN or inf", cnt[1], cnt[0]);
  }
  if (st) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, D.ev0, D.ev1) != hipSuccess) ms = 0.0f;
    memset(st, 0, sizeof *st);
    st->fallback_rows = cnt[2];
    st->short_rows = cnt[3];
    st->rowscan_rows = cnt[2];
    st->row_block = g_batch.QB;
    st->col_block = g_batch.TC;
    st->blocks_total = g_batch.blocks_total;
    st->candidates = g_batch.max_kp;
    st->col_splits = g_batch.max_split;
    st->path = 2;
    st->inserts = (uint64_t)cnt[4] | ((uint64_t)cnt[5] << 32);
    st->device_ms = ms;
    st->coarse_ms = ms;
    st->nonfinite_row = st->nonfinite_sample = UINT64_MAX;
    st->coarse_err_max = slots_max(cnt + 64);
    st->coarse_eps = g_batch.eps_list;
  }
  return 0;
}

}  // namespace

static int reweight_locked(Device& D, uint64_t n, uint64_t k, const uint32_t* idx, const double* cos,
                           const double* norms, uint32_t weight, double* out, hipStream_t s);

// ==================================================================================================
// C ABI

extern "C" {

const char* costal_last_error(void) { return g_err.c_str(); }

const char* costal_version(void) { return COSTAL_VERSION; }

int costal_device_info(char* buf, size_t len) {
  Device* Dp = call_device();
  if (!Dp) return COSTAL_E_DEVICE;
  Device& D = *Dp;
  std::lock_guard<std::mutex> lk(D.mu);
  int rc = device_init(D);
  if (rc) return rc;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, D.id));
  snprintf(buf, len, "%s (%s), %d CUs, %.1f GiB, clock %d kHz", prop.name, prop.gcnArchName,
           prop.multiProcessorCount, prop.totalGlobalMem / 1073741824.0, prop.clockRate);
  return 0;
}

int costal_knn_plan(uint64_t n, uint64_t d, const costal_knn_params* params, costal_knn_stats* stats) {
  if (!params || !stats) return fail(COSTAL_E_INVALID, "null pointer argument");
  int rc = check_args(n, d, params->k);
  if (rc) return rc;
  if (params->precision > COSTAL_PREC_FP16X3) return fail(COSTAL_E_INVALID, "unknown precision %u", params->precision);
  if (params->select > COSTAL_SELECT_TANIMOTO) return fail(COSTAL_E_INVALID, "unknown select %u", params->select);
  Device probe;  // the CU count of the current device when one is present, else 256
  probe.cus = 256;
  int dev = 0, count = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceCount(&count) == hipSuccess && count > 0 && hipGetDevice(&dev) == hipSuccess &&
      hipGetDeviceProperties(&prop, dev) == hipSuccess)
    probe.cus = prop.multiProcessorCount;
  (void)hipGetLastError();
  const Plan P = make_plan(probe, n, d, params->k, *params);
  memset(stats, 0, sizeof *stats);
  stats->nonfinite_row = stats->nonfinite_sample = UINT64_MAX;
  stats->path = P.fused ? 0 : P.wide ? 2 : 1;
  stats->row_block = (P.fused || P.wide) ? P.QB : 1;
  stats->col_block = (P.fused || P.wide) ? P.TC : 256;
  stats->candidates = P.kp;
  stats->col_splits = (P.fused || P.wide) ? P.nsplit : 1;
  stats->blocks_total = P.prune ? (uint64_t)P.order.nblk * P.order.npt
                       : (P.fused || P.wide) ? (uint64_t)((n + P.QB - 1) / P.QB) * ((n + P.TC - 1) / P.TC) : n;
  stats->coarse_eps = P.fused ? (P.tani ? tani_eps(fp16x3_eps((uint32_t)d))
                                       : P.prec == PREC_F16X3 ? fp16x3_eps((uint32_t)d) : fp32_eps((uint32_t)d))
                              : P.wide ? fp32_eps((uint32_t)d) : 0.0;
  return 0;
}

int costal_knn_exact_device(const float* x, uint64_t n, uint64_t d, const costal_knn_params* params,
                            uint32_t* idx, double* cos, double* norms, double* weights,
                            costal_knn_stats* stats, void* stream) {
  if (!params || !x || !idx || !cos) return fail(COSTAL_E_INVALID, "null pointer argument");
  int rc = check_args(n, d, params->k);
  if (rc) return rc;
  Device* Dp = call_device();
  if (!Dp) return COSTAL_E_DEVICE;
  Device& D = *Dp;
  std::lock_guard<std::mutex> lk(D.mu);
  rc = device_init(D);
  if (rc) return rc;
  return knn_device_locked(D, x, n, d, *params, idx, cos, norms, weights, stats,
                           pick_stream(D, stream), false);
}

int costal_mfma_probe(float* out) {
  if (!out) return fail(COSTAL_E_INVALID, "null pointer argument");
  Device* Dp = call_device();
  if (!Dp) return COSTAL_E_DEVICE;
  Device& D = *Dp;
  std::lock_guard<std::mutex> lk(D.mu);
  int rc = device_init(D);
  if (rc) return rc;
  if ((rc = ws_reserve(D, 256))) return rc;
  float* dout = reinterpret_cast<float*>(D.ws);
  mfma_probe_kernel<<<1, 64, 0, D.stream>>>(dout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, dout, 8 * sizeof(float), hipMemcpyDeviceToHost, D.stream));
  HIP_TRY(hipStreamSynchronize(D.stream));
  return 0;
}

int costal_knn_exact_device_noalloc(const float* x, uint64_t n, uint64_t d, const costal_knn_params* params,
                                    uint32_t* idx, double* cos, double* norms, double* weights,
                                    costal_knn_stats* stats, void* stream) {
  if (!params || !x || !idx || !cos) return fail(COSTAL_E_INVALID, "null pointer argument");
  int rc = check_args(n, d, params->k);
  if (rc) return rc;
  Device* Dp = call_device();
  if (!Dp) return COSTAL_E_DEVICE;
  Device& D = *Dp;
  std::lock_guard<std::mutex> lk(D.mu);
  rc = device_init(D);
  if (rc) return rc;
  if (!D.scratch)
    return fail(COSTAL_E_DEVICE, "the static scratch of device %d is not addressable: %s", D.id,
                hipGetErrorString(D.scratch_err));
  const FixedWs ws{D.scratch, (size_t)COSTAL_SCRATCH_BYTES};
  return knn_device_locked(D, x, n, d, *params, idx, cos, norms, weights, stats, pick_stream(D, stream), false, &ws);
}

static int knn_host(const float* x, uint64_t n, uint64_t d, const costal_knn_params* params, uint32_t* idx,
                    double* cos, double* norms, double* weights, costal_knn_stats* stats, bool brute) {
  if (!params || !x || !idx || !cos) return fail(COSTAL_E_INVALID, "null pointer argument");
  int rc = check_args(n, d, params->k);
  if (rc) return rc;
  Device* Dp = call_device();
  if (!Dp) return COSTAL_E_DEVICE;
  Device& D = *Dp;
  std::lock_guard<std::mutex> lk(D.mu);
  rc = device_init(D);
  if (rc) return rc;
  const uint64_t k = params->k;
  hipStream_t s = D.stream;
  const bool mlp = params->mlp_params && !params->row_threshold;
  const size_t mlp_count = mlp ? costal_mlp_param_count(d) : 0;
  // the call's device copies, carved from the grow-only buffer
  auto carve = [&](Carve& c, float*& dx, uint32_t*& didx, double*& dcos, double*& dnorms, double*& dw, float*& dth,
                   float*& dmlp) {
    dx = c.take<float>(std::max<size_t>(n * d, 1));
    didx = c.take<uint32_t>(n * k);
    dcos = c.take<double>(n * k);
    dnorms = c.take<double>(n);
    dw = weights ? c.take<double>(n * k) : nullptr;
    dth = params->row_threshold ? c.take<float>(n) : nullptr;
    dmlp = mlp ? c.take<float>(mlp_count) : nullptr;
  };
  float *dx, *dth, *dmlp;
  uint32_t* didx;
  double *dcos, *dnorms, *dw;
  {
    Carve c(nullptr);
    carve(c, dx, didx, dcos, dnorms, dw, dth, dmlp);
    rc = io_reserve(D, c.off + 256);
    if (rc) return rc;
  }
  Carve c(D.io);
  carve(c, dx, didx, dcos, dnorms, dw, dth, dmlp);
  costal_knn_params p = *params;
  if (mlp) {
    HIP_TRY(hipMemcpyAsync(dmlp, params->mlp_params, mlp_count * sizeof(float), hipMemcpyHostToDevice, s));
    p.mlp_params = dmlp;
  }
  if (dth) {
    HIP_TRY(hipMemcpyAsync(dth, params->row_threshold, n * sizeof(float), hipMemcpyHostToDevice, s));
    p.row_threshold = dth;
  }
  HIP_TRY(hipMemcpyAsync(dx, x, n * d * sizeof(float), hipMemcpyHostToDevice, s));
  rc = knn_device_locked(D, dx, n, d, p, didx, dcos, dnorms, dw, stats, s, brute);
  if (rc) {
    (void)hipStreamSynchronize(s);
    return rc;
  }
  HIP_TRY(hipMemcpyAsync(idx, didx, n * k * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(cos, dcos, n * k * sizeof(double), hipMemcpyDeviceToHost, s));
  if (norms) HIP_TRY(hipMemcpyAsync(norms, dnorms, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (weights) HIP_TRY(hipMemcpyAsync(weights, dw, n * k * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int costal_knn_exact(const float* x, uint64_t n, uint64_t d, const costal_knn_params* params,
                     uint32_t* idx, double* cos, double* norms, double* weights,
                     costal_knn_stats* stats) {
  return knn_host(x, n, d, params, idx, cos, norms, weights, stats, false);
}

int costal_knn_batched_device(const costal_sample* samples, uint32_t nsamples, const costal_knn_params* params,
                              costal_knn_stats* stats, void* stream) {
  if (!params || (!samples && nsamples)) return fail(COSTAL_E_INVALID, "null pointer argument");
  if (nsamples == 0) return 0;
  Device* Dp = call_device();
  if (!Dp) return COSTAL_E_DEVICE;
  Device& D = *Dp;
  std::lock_guard<std::mutex> lk(D.mu);
  int rc = device_init(D);
  if (rc) return rc;
  return knn_batched_locked(D, samples, nsamples, *params, stats, pick_stream(D, stream));
}

int costal_knn_bruteforce(const float* x, uint64_t n, uint64_t d, uint64_t k, uint32_t select,
                          uint32_t* idx, double* cos, double* norms) {
  costal_knn_params p;
  memset(&p, 0, sizeof p);
  p.k = k;
  p.margin = -1;
  p.select = select;
  return knn_host(x, n, d, &p, idx, cos, norms, nullptr, nullptr, true);
}

int costal_reweight_device(uint64_t n, uint64_t k, const uint32_t* idx, const double* cos,
                           const double* norms, uint32_t weight, double* out, void* stream) {
  if (weight != COSTAL_WEIGHT_TANIMOTO && weight != COSTAL_WEIGHT_COSINE)
    return fail(COSTAL_E_INVALID, "weight %u is not computed on the device", weight);
  Device* Dp = call_device();
  if (!Dp) return COSTAL_E_DEVICE;
  Device& D = *Dp;
  std::lock_guard<std::mutex> lk(D.mu);
  int rc = device_init(D);
  if (rc) return rc;
  return reweight_locked(D, n, k, idx, cos, norms, weight, out, pick_stream(D, stream));
}

// weight is Tanimoto or cosine; D.mu held, device initialised
static int reweight_locked(Device& D, uint64_t n, uint64_t k, const uint32_t* idx, const double* cos,
                           const double* norms, uint32_t weight, double* out, hipStream_t s) {
  if (n * k == 0) return 0;
  int rc = ws_reserve(D, 256);
  if (rc) return rc;
  uint32_t* bad = (uint32_t*)D.ws;
  uint32_t init = NONE, got = NONE;
  HIP_TRY(hipMemcpyAsync(bad, &init, sizeof init, hipMemcpyHostToDevice, s));
  const uint64_t total = n * k;
  reweight_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(n, (uint32_t)k, idx, cos, norms, weight, out,
                                                                   bad);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&got, bad, sizeof got, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (got != NONE) return fail(COSTAL_E_INVALID, "neighbor index out of range (n=%llu)", (unsigned long long)n);
  return 0;
}

int costal_reweight(uint64_t n, uint64_t k, const uint32_t* idx, const double* cos, const double* norms,
                    uint32_t weight, double* out) {
  if (!idx || !cos || !norms || !out) return fail(COSTAL_E_INVALID, "null pointer argument");
  for (uint64_t e = 0; e < n * k; ++e)
    if (idx[e] >= n)
      return fail(COSTAL_E_INVALID, "neighbor index %u out of range (n=%llu)", idx[e], (unsigned long long)n);
  if (weight == COSTAL_WEIGHT_JACCARD) {
    // weight.rs:64-93: O(N k^2) set intersection of sorted neighbor lists (integer work on the host)
    std::vector<uint32_t> sorted(idx, idx + n * k);
    for (uint64_t i = 0; i < n; ++i) std::sort(sorted.begin() + i * k, sorted.begin() + (i + 1) * k);
    for (uint64_t i = 0; i < n; ++i) {
      const uint32_t* a = sorted.data() + i * k;
      for (uint64_t t = 0; t < k; ++t) {
        const uint32_t* b = sorted.data() + (uint64_t)idx[i * k + t] * k;
        uint64_t p = 0, q = 0, sh = 0;
        while (p < k && q < k) {
          if (a[p] < b[q]) ++p;
          else if (a[p] > b[q]) ++q;
          else { ++sh; ++p; ++q; }
        }
        out[i * k + t] = (double)sh / (2.0 * (double)k - (double)sh);
      }
    }
    return 0;
  }
  if (weight != COSTAL_WEIGHT_TANIMOTO && weight != COSTAL_WEIGHT_COSINE)
    return fail(COSTAL_E_INVALID, "unknown edge weight %u", weight);
  if (n * k == 0) return 0;
  Device* Dp = call_device();
  if (!Dp) return COSTAL_E_DEVICE;
  Device& D = *Dp;
  std::lock_guard<std::mutex> lk(D.mu);
  int rc = device_init(D);
  if (rc) return rc;
  hipStream_t s = D.stream;
  {
    Carve c(nullptr);
    c.take<uint32_t>(n * k);
    c.take<double>(n * k);
    c.take<double>(n);
    c.take<double>(n * k);
    rc = io_reserve(D, c.off + 256);
    if (rc) return rc;
  }
  Carve c(D.io);
  uint32_t* didx = c.take<uint32_t>(n * k);
  double* dcos = c.take<double>(n * k);
  double* dn = c.take<double>(n);
  double* dout = c.take<double>(n * k);
  HIP_TRY(hipMemcpyAsync(didx, idx, n * k * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dcos, cos, n * k * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dn, norms, n * 8, hipMemcpyHostToDevice, s));
  rc = reweight_locked(D, n, k, didx, dcos, dn, weight, dout, s);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, dout, n * k * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int costal_csr_from_edges(uint64_t n, uint64_t m, const uint32_t* rows, const uint32_t* cols, const double* w,
                          uint64_t* indptr, uint32_t* nbr, double* wout, uint64_t* nnz) {
  // leiden.rs:52-100: validate in edge order (range, self-loop, weight), count the kept entries
  // of both endpoints, prefix-sum the counts into row starts, then place each kept edge in both
  // rows in edge order.
  if (!indptr || !nnz || (m && (!rows || !cols || !w || !nbr || !wout)))
    return fail(COSTAL_E_INVALID, "null pointer argument");
  if (n > 0xffffffffull) return fail(COSTAL_E_TOO_MANY_ROWS, "%llu nodes exceed the u32 node index", (unsigned long long)n);
  std::vector<uint64_t> deg(n + 1, 0);
  for (uint64_t e = 0; e < m; ++e) {
    const uint64_t a = rows[e], b = cols[e];
    if (a >= n || b >= n)
      return fail(COSTAL_E_INVALID, "edge %llu (%llu, %llu) is out of range (n=%llu)", (unsigned long long)e,
                  (unsigned long long)a, (unsigned long long)b, (unsigned long long)n);
    if (a == b) return fail(COSTAL_E_INVALID, "edge %llu is a self-loop on node %llu", (unsigned long long)e,
                            (unsigned long long)a);
    const double x = w[e];
    if (!std::isfinite(x) || x < 0.0)
      return fail(COSTAL_E_INVALID, "edge %llu has weight %s; weights must be finite and non-negative",
                  (unsigned long long)e, shortest(x).c_str());
    if (x != 0.0) {
      ++deg[a + 1];
      ++deg[b + 1];
    }
  }
  for (uint64_t v = 1; v <= n; ++v) deg[v] += deg[v - 1];
  std::copy(deg.begin(), deg.end(), indptr);
  for (uint64_t e = 0; e < m; ++e) {
    if (w[e] == 0.0) continue;
    const uint64_t pa = deg[rows[e]]++, pb = deg[cols[e]]++;
    nbr[pa] = cols[e];
    wout[pa] = w[e];
    nbr[pb] = rows[e];
    wout[pb] = w[e];
  }
  *nnz = indptr[n];
  return 0;
}

int costal_fuzzy_union(uint64_t n, uint64_t k, const uint32_t* idx, const double* w, uint32_t* rows,
                       uint32_t* cols, double* wout, uint64_t* m) {
  // graph.rs:24-68: key = lo << 32 | hi, sort, combine (a + b) - a*b, drop exact zeros.
  if (!idx || !w || !rows || !cols || !wout || !m) return fail(COSTAL_E_INVALID, "null pointer argument");
  for (uint64_t e = 0; e < n * k; ++e)
    if (idx[e] >= n)
      return fail(COSTAL_E_INVALID, "neighbor index %u out of range (n=%llu)", idx[e], (unsigned long long)n);
  std::vector<std::pair<uint64_t, double>> e;
  e.reserve(n * k);
  for (uint64_t t = 0; t < n * k; ++t) {
    const uint64_t i = t / k, j = idx[t];
    if (i == j) continue;
    const uint64_t lo = i < j ? i : j, hi = i < j ? j : i;
    e.emplace_back((lo << 32) | hi, w[t]);
  }
  std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  uint64_t out = 0;
  for (size_t t = 0; t < e.size();) {
    const uint64_t key = e[t].first;
    const double a = e[t].second;
    double v;
    if (t + 1 < e.size() && e[t + 1].first == key) {
      const double b = e[t + 1].second;
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
  *m = out;
  return 0;
}

}  // extern "C"
