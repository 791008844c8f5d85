// scan.h -- the fused MFMA tile + candidate-list scan (knn.rs:603-660, 759-875), shared by the
// instantiation units (scan_inst.hip, one object per precision form and list capacity) and the
// host code that launches it.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "../../include/costal_hip.h"
#include "device_ids.h"

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

namespace costal {

struct BatchSample {
  const float* x;        // the sample's rows (device), n x d
  uint32_t n, d;
  uint32_t row_pad;      // first row of the sample in the stacked working copies
  uint32_t row_base;     // first row of the sample in the concatenated row numbering
  int kp;                // candidate-list length K'
  uint32_t nsplit, cols_per_split;
  uint64_t slot_base;    // first candidate list (slot) of the sample
  uint64_t entry_base;   // first list entry of the sample (its lists hold kp entries each)
  uint32_t* idx;
  double* cos;
  double* norms;
  double* wts;
};

typedef void (*CoarseFn)(const uint8_t*, const float*, uint32_t, uint32_t, int, const uint32_t*,
                         uint32_t, const float*, uint32_t, uint32_t, float, uint32_t*, float*, uint32_t*,
                         float*, unsigned long long*);
typedef void (*CoarseBatchFn)(const BatchSample*, const uint32_t*, uint32_t, uint32_t*, const uint8_t*,
                              const float*, uint32_t, float, uint32_t*, float*, uint32_t*, float*,
                              unsigned long long*, const unsigned long long*);

// One instantiation unit per (precision form, list capacity): the kernel for J 16-byte row chunks
// per lane half, or null when the form has none for that J.
#define COSTAL_SCAN_DECL(P, K)                       \
  CoarseFn scan_fn_p##P##_k##K(uint32_t J);          \
  CoarseBatchFn scan_batch_fn_p##P##_k##K(uint32_t J);
COSTAL_SCAN_DECL(0, 16) COSTAL_SCAN_DECL(0, 32) COSTAL_SCAN_DECL(0, 64)
COSTAL_SCAN_DECL(1, 16) COSTAL_SCAN_DECL(1, 32) COSTAL_SCAN_DECL(1, 64)
COSTAL_SCAN_DECL(2, 16) COSTAL_SCAN_DECL(2, 32) COSTAL_SCAN_DECL(2, 64)
#undef COSTAL_SCAN_DECL

}  // namespace costal

namespace {

using costal::BatchSample;

constexpr int PREC_F32 = COSTAL_PREC_FP32;
constexpr int PREC_F16 = COSTAL_PREC_FP16;
constexpr uint32_t NONE = 0xffffffffu;

// f32 inner product of two zero-padded normalized rows, sequential fma: |err| <= gamma_d.
__device__ __forceinline__ float dot32(const float* __restrict__ a, const float* __restrict__ b,
                                       uint32_t dp) {
  float s = 0.0f;
  const f32x4* a4 = reinterpret_cast<const f32x4*>(a);
  const f32x4* b4 = reinterpret_cast<const f32x4*>(b);
  for (uint32_t t = 0; t < dp / 4; ++t) {
    f32x4 u = a4[t], v = b4[t];
    s = __builtin_fmaf(u[0], v[0], s);
    s = __builtin_fmaf(u[1], v[1], s);
    s = __builtin_fmaf(u[2], v[2], s);
    s = __builtin_fmaf(u[3], v[3], s);
  }
  return s;
}

// ------------------------------------------------------------------------------------------------
// coarse_kernel
//
// Workgroup = NW waves; each wave owns 64 queries in two groups of 32, held in registers for the
// whole launch.  The workgroup streams its candidate range through two LDS buffers of TC-row
// tiles (about 24 KB each), copied global -> LDS by the load units directly (global_load_lds,
// 16 bytes per lane) one tile ahead of the compute, so no registers hold staged data.  For every
// 32-candidate sub-tile a wave issues the MFMAs D[cand][query] = sum_k cand[k] * query[k] for both
// query groups: in each 32x32 accumulator, lane l holds query (l % 32) of the group and the 16
// candidates 8*(i/4) + 4*(l/32) + (i%4), i = 0..15.  Lane l owns query l of the wave (group l / 32,
// column l % 32) and keeps its one sorted candidate list in registers; the other half of its
// query's scores sits in lane l ^ 32 (which owns a query of the other group).  The fast path
// compares the maxima of both accumulators with the thresholds of the queries they belong to
// (the lane's own, and the partner lane's, exchanged once per tile; a stale partner threshold is
// only lower, so it never turns a candidate away).  A wave with a hit exchanges the accumulator
// halves with v_permlane32_swap, so each lane holds all 32 scores of its own query, and inserts.
// MFMAs run one step ahead of the scan (see scan_block).
//
// Row layout (global and LDS): a row is 2*J 16-byte chunks; chunk q feeds lane half h = q % 2,
// slot j = q / 2.  Queries and candidates use the same chunk -> (lane half, slot) map, so each
// MFMA contracts matching feature indices (any consistent permutation of k leaves the sum a sum of
// the same products).  In LDS a tile is 1 KB groups, group (sub-tile s, slot j) = s*J + j holding
// the two chunks 2j, 2j+1 of the sub-tile's 32 rows, lane-contiguous (slot 32*half + row): each
// ds_read_b128 is conflict free, and each group is one wave-wide global_load_lds (lane = slot).
__device__ __forceinline__ float max16(const f32x16& a) {
  const float m0 = fmaxf(fmaxf(a[0], a[1]), a[2]);
  const float m1 = fmaxf(fmaxf(a[3], a[4]), a[5]);
  const float m2 = fmaxf(fmaxf(a[6], a[7]), a[8]);
  const float m3 = fmaxf(fmaxf(a[9], a[10]), a[11]);
  const float m4 = fmaxf(fmaxf(a[12], a[13]), a[14]);
  return fmaxf(fmaxf(fmaxf(m0, m1), m2), fmaxf(fmaxf(m3, m4), a[15]));
}

// a[i] for a lane-varying i, as a select tree on the bits of i (keeps the accumulator in registers)
__device__ __forceinline__ float pick16(const f32x16& a, int i) {
  const bool b3 = i & 8, b2 = i & 4, b1 = i & 2, b0 = i & 1;
  const float c0 = b3 ? a[8] : a[0], c1 = b3 ? a[9] : a[1], c2 = b3 ? a[10] : a[2], c3 = b3 ? a[11] : a[3];
  const float c4 = b3 ? a[12] : a[4], c5 = b3 ? a[13] : a[5], c6 = b3 ? a[14] : a[6], c7 = b3 ? a[15] : a[7];
  const float d0 = b2 ? c4 : c0, d1 = b2 ? c5 : c1, d2 = b2 ? c6 : c2, d3 = b2 ? c7 : c3;
  const float e0 = b1 ? d2 : d0, e1 = b1 ? d3 : d1;
  return b0 ? e1 : e0;
}

// The value held by the lane of the other half of the wave (lane l ^ 32): the other list of the
// same query.
__device__ __forceinline__ float partner32(float x) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float((__builtin_amdgcn_workitem_id_x() & 32) ? r[0] : r[1]);
}

// Sorted (ascending) register list: ls[0] is the smallest score.  Entries [kp, KPR) hold +inf and
// never move; empty entries hold the admission threshold.  push() replaces the minimum and bubbles
// the new entry up to its place.
template <int KPR>
__device__ __forceinline__ void list_push(float (&ls)[KPR], uint32_t (&li)[KPR], float v, uint32_t c) {
  ls[0] = v;
  li[0] = c;
#pragma unroll
  for (int e = 0; e < KPR - 1; ++e) {
    const float a = ls[e], b = ls[e + 1];
    if (!(a > b)) break;
    const uint32_t ia = li[e], ib = li[e + 1];
    ls[e] = b;
    ls[e + 1] = a;
    li[e] = ib;
    li[e + 1] = ia;
  }
}

// Precision forms of the coarse tile:
//   PREC_F32   f32 rows, v_mfma_f32_32x32x2f32; a 16-byte chunk is 4 features (4 MFMAs).
//   PREC_F16   FP16 rows, v_mfma_f32_32x32x16_f16; a chunk is 8 features (1 MFMA); a hit is re-scored
//              in f32 before admission.
//   PREC_F16X3 FP16 hi/lo split rows (x = hi + lo + O(2^-22)): chunks [0, J/2) hold hi, [J/2, J) lo;
//              hi.hi + hi.lo + lo.hi per chunk pair (3 MFMAs), an f32-class score without re-scoring.
constexpr int PREC_F16X3 = COSTAL_PREC_FP16X3;

// The 32x32 product of one sub-tile: b = the lane's J candidate chunks, q = its J query chunks.
template <int PREC, int J>
__device__ __forceinline__ f32x16 tile_mfma(const uint4 (&b)[J], const uint4 (&q)[J], f32x16 acc = {}) {
  if constexpr (PREC == PREC_F32) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const f32x4 b4 = __builtin_bit_cast(f32x4, b[j]);
      const f32x4 q4 = __builtin_bit_cast(f32x4, q[j]);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(b4[0], q4[0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(b4[1], q4[1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(b4[2], q4[2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(b4[3], q4[3], acc, 0, 0, 0);
    }
  } else if constexpr (PREC == PREC_F16) {
#pragma unroll
    for (int j = 0; j < J; ++j)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, b[j]), __builtin_bit_cast(f16x8, q[j]),
                                                   acc, 0, 0, 0);
  } else {
    constexpr int JH = J / 2;
#pragma unroll
    for (int j = 0; j < JH; ++j) {
      const f16x8 ch = __builtin_bit_cast(f16x8, b[j]), cl = __builtin_bit_cast(f16x8, b[j + JH]);
      const f16x8 qh = __builtin_bit_cast(f16x8, q[j]), ql = __builtin_bit_cast(f16x8, q[j + JH]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(cl, qh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ch, ql, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ch, qh, acc, 0, 0, 0);
    }
  }
  return acc;
}

template <int J>
struct TileShape {
  static constexpr int ROWB = 32 * J;  // bytes per working-copy row
  // candidates per tile: about 24 KB tiles, so three four-wave workgroups fit a CU's LDS (two
  // for J >= 7); an even number of 32-candidate steps, at least four (the pipeline reads three
  // steps ahead)
  static constexpr int TC = J <= 1 ? 768 : J <= 2 ? 384 : J <= 3 ? 256 : J <= 4 ? 192 : 128;
  static constexpr int TILEB = TC * ROWB;
  static constexpr int NSTEP = TC / 32;      // one 32-candidate sub-tile per scan step
  static constexpr int GROUPS = TILEB / 1024;  // 1 KB LDS groups of a tile
  static_assert(NSTEP % 2 == 0 && NSTEP >= 4, "pipeline steps per tile");
};

// Rows of zero padding every working copy carries past n: a whole number of tiles of every shape
// (768, 384, 256, 192, 128 rows) and of the wide-row blocks (256 rows).
constexpr uint32_t PAD_ROWS = 1536;


// One work item of the scan: query block `qblock` (QB = 64 NW queries) against candidates
// [c_begin, c_end) of one sample; the lists go to split `split` of the queries' nsplit lists.
// `tiles` is the workgroup's two LDS tile buffers.
//
// Software pipeline over scan steps (one 32-candidate sub-tile each, NSTEP per tile): step p first
// reads the candidate chunks of step p+2 from LDS into the chunk registers the MFMAs of step p-1
// have consumed, then issues the MFMAs of step p+1 and scans the accumulators of step p, the MFMAs
// interleaved with the scan's compares so that the wave never waits on the matrix core to issue
// them; the LDS reads have a whole step to complete.  Accumulators and chunk registers are double
// buffered by the parity of p (NSTEP is even).  At step NSTEP-2 of tile t every wave waits for its
// share of tile t+1's copy, the workgroup meets, and the copy of tile t+2 into the buffer tile t
// leaves is issued: it has a whole tile of steps to arrive.
template <int PREC, int J, int KPR, int NW>
__device__ __forceinline__ void scan_block(
    uint8_t* __restrict__ tiles,       // 2 * TILEB bytes of LDS
    const uint8_t* __restrict__ xsq,   // n_pad rows of 32*J bytes (the working copy of this form)
    const float* __restrict__ xs32,    // f32 working copy (dp32 per row), for the FP16 re-score
    uint32_t dp32, uint32_t n, int kp,
    const uint32_t* __restrict__ rowlist, uint32_t nq,  // queries: rowlist[0..nq) or rows 0..nq-1
    uint32_t qblock, const float* __restrict__ tinit,   // per-row admission threshold (by row) or null
    uint32_t split, uint32_t nsplit,
    uint32_t c_begin, uint32_t c_end,  // candidate range; c_begin a multiple of the tile, c_end <= n
    float eps_pre,                     // FP16 form: bound on |FP16 score - f32 score|
    uint32_t* __restrict__ cand_idx, float* __restrict__ cand_sc, uint32_t* __restrict__ cand_cnt,
    float* __restrict__ cand_floor, unsigned long long* __restrict__ inserts) {
  using TS = TileShape<J>;
  constexpr int QB = 64 * NW;
  constexpr int TC = TS::TC, ROWB = TS::ROWB, NSTEP = TS::NSTEP, TILEB = TS::TILEB;
  constexpr int RD = 2;  // chunks of step p + RD are read at step p
  static_assert(TS::GROUPS % NW == 0, "every wave copies the same number of groups");
  static_assert(RD == 2, "the tile change is the step before the last");
  typedef __attribute__((address_space(3))) void lds_t;
  typedef __attribute__((address_space(1))) const void glob_t;

  const int tid = tid_x();
  const int l = tid & 63, h = l >> 5, l32 = l & 31;
  const uint32_t w = __builtin_amdgcn_readfirstlane((uint32_t)tid >> 6);
  const uint32_t qbase = qblock * QB + w * 64;
  auto row_of = [&](uint32_t q) -> uint32_t { return q < nq ? (rowlist ? rowlist[q] : q) : 0u; };
  // the two query groups' rows, as B operands: lane l gives column l % 32 of each group
  const uint32_t g0 = row_of(qbase + l32), g1 = row_of(qbase + 32 + l32);
  uint4 q0[J], q1[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    q0[j] = *reinterpret_cast<const uint4*>(xsq + (size_t)g0 * ROWB + (2 * j + h) * 16);
    q1[j] = *reinterpret_cast<const uint4*>(xsq + (size_t)g1 * ROWB + (2 * j + h) * 16);
  }
  // the lane's own query: number l of the wave
  const uint32_t qq = qbase + l;
  const bool qvalid = qq < nq;
  const uint32_t qg = h ? g1 : g0;
  // Empty list entries hold the lane's admission threshold (-inf without one; +inf for padding
  // queries, which so never insert): the list minimum ls[0] is then always the lane's threshold,
  // and real entries (score above it) are the ones with an index.  A NaN threshold (caller data
  // or a model with NaN parameters) admits everything: the test reads the bits, since this code is
  // built without NaN semantics.
  float t0 = qvalid ? (tinit ? tinit[qg] : -INFINITY) : INFINITY;
  if ((__float_as_uint(t0) & 0x7fffffffu) > 0x7f800000u) t0 = -INFINITY;
  float ls[KPR];
  uint32_t li[KPR];
#pragma unroll
  for (int e = 0; e < KPR; ++e) {
    ls[e] = e < kp ? t0 : INFINITY;
    li[e] = NONE;
  }
  float th = ls[0];
  float thp = partner32(th);  // the partner lane's threshold, refreshed once per tile
  // fast-path thresholds of the two accumulators (group 0, group 1), minus the FP16 pre-filter
  // bound in that form
  const float pre = PREC == PREC_F16 ? eps_pre : 0.0f;
  float fA = (h ? thp : th) - pre, fB = (h ? th : thp) - pre;
  uint32_t my_inserts = 0;
  const float* xq32 = xs32 + (size_t)qg * dp32;

  const int ntile = c_begin < c_end ? (int)((c_end - c_begin + TC - 1) / TC) : 0;
  // tile T of the range into LDS buffer b: wave w copies groups w, w + NW, ...; group g = s*J + j
  // is the chunks 2j, 2j+1 of rows 32s..32s+31, lane l taking chunk 2j + l/32 of row 32s + l%32
  // (the wave-uniform part of each address first, the lane's 32-bit offset last, so a load can
  // take a scalar base and one vector offset)
  const uint32_t lane_off = (uint32_t)l32 * ROWB + 16 * h;
  auto copy_tile = [&](int T, uint8_t* dst) __attribute__((always_inline)) {
    const uint8_t* src = xsq + (size_t)(c_begin + (uint32_t)T * TC) * ROWB;
#pragma unroll
    for (int c = 0; c < TS::GROUPS / NW; ++c) {
      const uint32_t g = w + NW * c;
      const uint32_t s = g / J, j = g % J;
      const uint8_t* base = src + (size_t)(s * 32 * ROWB + 32 * j);
      uint32_t lo = lane_off;
      __asm__ volatile("" : "+v"(lo));  // recomputed per copy, not a hoisted 64-bit address
      __builtin_amdgcn_global_load_lds((glob_t*)(base + lo), (lds_t*)(dst + g * 1024), 16, 0, 0);
    }
  };
  // this wave's copies complete, every wave's LDS reads complete, then the workgroup meets
  auto tile_sync = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  };

  f32x16 acc[2][2];  // [step parity][query group]
  uint4 bop[2][J];   // [step parity][slot]: the lane's candidate chunks of one sub-tile
  auto read_b = [&](int slot, const uint8_t* buf, int s) __attribute__((always_inline)) {
    const uint8_t* p = buf + s * J * 1024 + l * 16;
#pragma unroll
    for (int j = 0; j < J; ++j) bop[slot][j] = *reinterpret_cast<const uint4*>(p + j * 1024);
  };
  auto mfma_step = [&](int slot) __attribute__((always_inline)) {
    acc[slot][0] = tile_mfma<PREC, J>(bop[slot], q0);
    acc[slot][1] = tile_mfma<PREC, J>(bop[slot], q1);
  };
  // MFMAs per step: both query groups' tiles
  constexpr int NMFMA = 2 * (PREC == PREC_F32 ? 4 * J : PREC == PREC_F16 ? J : 3 * (J / 2));
  // one candidate of the lane's own query: admitted when it beats the threshold
  auto offer = [&](float v, uint32_t cand) __attribute__((always_inline)) {
    if (v > th - pre && cand < c_end && cand != qg) {
      if constexpr (PREC == PREC_F16) {
        v = dot32(xq32, xs32 + (size_t)cand * dp32, dp32);
        if (!(v > th)) return;
      }
      list_push<KPR>(ls, li, v, cand);
      th = ls[0];
      ++my_inserts;
    }
  };
  // scan the accumulators of one step whose candidates start at cbase.  Fast path: the maxima of
  // both accumulators against the thresholds of the queries they belong to.  Slow path: the
  // quads (elements 4q..4q+3: candidates 8q + 4h + 0..3) where some lane has a hit (a
  // wave-uniform mask), and in each such quad every element exchanged with the partner lane, so
  // the lane offers both halves of its own query's scores (candidates 8q + e and 8q + e + 4).
  auto fast_test = [&](const f32x16& a, const f32x16& b) __attribute__((always_inline)) -> bool {
    const float ma = max16(a), mb = max16(b);
    // two lane masks or-ed in scalar registers (no short-circuit branch between the maxima)
    return (__builtin_amdgcn_ballot_w64(ma > fA) | __builtin_amdgcn_ballot_w64(mb > fB)) != 0;
  };
  auto slow_path = [&](f32x16& a, f32x16& b, uint32_t cbase) __attribute__((always_inline)) {
    uint32_t qmask = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float qa = fmaxf(fmaxf(fmaxf(a[4 * q], a[4 * q + 1]), a[4 * q + 2]), a[4 * q + 3]);
      const float qb = fmaxf(fmaxf(fmaxf(b[4 * q], b[4 * q + 1]), b[4 * q + 2]), b[4 * q + 3]);
      qmask |= (__builtin_amdgcn_ballot_w64(qa > fA) | __builtin_amdgcn_ballot_w64(qb > fB)) != 0 ? 1u << q : 0u;
    }
    while (qmask) {
      const int q = __builtin_ctz(qmask);
      qmask &= qmask - 1;
#pragma unroll 1
      for (int e = 0; e < 4; ++e) {
        const int i = 4 * q + e;
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a[i]), __float_as_uint(b[i]), false, false);
        const uint32_t c = cbase + 8 * q + e;
        offer(__uint_as_float(r[0]), c);
        offer(__uint_as_float(r[1]), c + 4);
      }
    }
    fA = (h ? thp : th) - pre;
    fB = (h ? th : thp) - pre;
  };
  // step p of tile t (parity S = p & 1): the tile change at step NSTEP-RD, then the chunks of step
  // p+RD, the MFMAs of step p+1 and the scan of step p.  The step opens with a wait for every LDS
  // read in flight, which are the chunks its MFMAs use (read a step earlier): the copies into LDS
  // leave the compiler no in-order count of LDS reads, so without this wait it would wait at the
  // first MFMA for the reads this step has just issued as well.
  int t = 0;
  auto step = [&](int p, int S, bool change) __attribute__((always_inline)) {
    if (p == 0) {
      thp = partner32(th);
      fA = (h ? thp : th) - pre;
      fB = (h ? th : thp) - pre;
    }
    const int cur = t & 1;
    if (change) {
      tile_sync();
      if (t + 2 < ntile) copy_tile(t + 2, tiles + cur * TILEB);
    } else {
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    }
    // one scheduling region: the chunks of step p+RD into bop[S] (read by the MFMAs of the step
    // before); the MFMAs of step p+1 (chunks in bop[1-S]); and the fast-path compares of step p,
    // the MFMAs spread over the compares
    const int src = p + RD < NSTEP ? cur : 1 - cur;
    const int sstep = p + RD < NSTEP ? p + RD : p + RD - NSTEP;
    read_b(S, tiles + src * TILEB, sstep);
    mfma_step(1 - S);
    const bool hit = fast_test(acc[S][0], acc[S][1]);
    __builtin_amdgcn_sched_group_barrier(0x100, J, 0);  // the LDS reads first
#pragma unroll
    for (int m = 0; m < NMFMA; ++m) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
      __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);  // then three VALU
    }
    if (hit) slow_path(acc[S][0], acc[S][1], c_begin + (uint32_t)t * TC + (uint32_t)p * 32);
  };

  if (ntile > 0) {
    copy_tile(0, tiles);
    tile_sync();
    if (ntile > 1) copy_tile(1, tiles + TILEB);
    read_b(0, tiles, 0);
    read_b(1, tiles, 1);
    mfma_step(0);
    // the tile change peeled from the step loop: its copies stay out of the other steps' paths
    for (;;) {
      for (int p = 0; p < NSTEP - RD; p += 2) {
        step(p, 0, false);
        step(p + 1, 1, false);
      }
      step(NSTEP - 2, 0, true);
      step(NSTEP - 1, 1, false);
      if (++t >= ntile) break;
    }
  }

  if (inserts) {
    // one atomic per wave
    uint32_t tot = my_inserts;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (l == 0) atomicAdd(inserts, (unsigned long long)tot);
  }
  if (qvalid) {
    const size_t slot = (size_t)qq * nsplit + split;
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
    // everything the list turned away scored at most this (the fast path compared the other
    // half of the query's scores with an older, lower copy of it)
    cand_floor[slot] = th;
  }
}

// Two launch shapes:
//   units_per_wg == 0: grid (query blocks, splits); workgroup (x, y) scans query block x against
//     columns [y * cols_per_split, (y+1) * cols_per_split).
//   units_per_wg > 0: one workgroup per resident slot, each taking the same number of (query block,
//     tile) units in query-block-major order, so every slot ends together.  A query block is cut
//     into the pieces that fall to different workgroups; piece p of a block (p = workgroup -
//     the first workgroup touching the block) fills list p of its queries, and the workgroup with
//     the block's last piece marks the lists after it empty.  cols_per_split then gives the lists
//     per query (the most pieces a block can have: ceil(tiles / units_per_wg) + 1).
// Waves per SIMD the kernel is built for: three where the registers allow it (narrow rows, short
// lists), two for the rest, and no bound for the longest lists on wide rows (which would spill
// to scratch under one; they serve rescans of a few rows).
template <int J, int KPR>
constexpr int scan_waves() {
  return (J <= 2 && KPR <= 16) ? 3 : (KPR >= 64 || (J >= 7 && KPR >= 32)) ? 1 : 2;
}

template <int PREC, int J, int KPR, int NW>
__global__ __launch_bounds__(64 * NW, (scan_waves<J, KPR>())) void coarse_kernel(
    const uint8_t* __restrict__ xsq, const float* __restrict__ xs32, uint32_t dp32, uint32_t n, int kp,
    const uint32_t* __restrict__ rowlist, uint32_t nq, const float* __restrict__ tinit, uint32_t cols_per_split,
    uint32_t units_per_wg, float eps_pre, uint32_t* __restrict__ cand_idx, float* __restrict__ cand_sc,
    uint32_t* __restrict__ cand_cnt, float* __restrict__ cand_floor, unsigned long long* __restrict__ inserts) {
  __shared__ __attribute__((aligned(1024))) uint8_t tiles[2 * TileShape<J>::TILEB];
  constexpr uint32_t TC = TileShape<J>::TC;
  constexpr uint32_t QB = 64 * NW;
  const uint32_t ntile = (n + TC - 1) / TC;
  const uint64_t units = units_per_wg ? (uint64_t)((nq + QB - 1) / QB) * ntile : 1;
  uint64_t u = units_per_wg ? (uint64_t)bid_x() * units_per_wg : 0;
  const uint64_t u_end = units_per_wg ? (u + units_per_wg < units ? u + units_per_wg : units) : 1;
  // one piece per iteration (the direct shape has exactly one); a single call site keeps one copy
  // of the scan in the kernel
  for (bool first = true; u < u_end; first = false) {
    uint32_t qb, split, nspl, c0, c1, len;
    bool whole = false;
    if (units_per_wg) {
      qb = (uint32_t)(u / ntile);
      const uint32_t tb = (uint32_t)(u % ntile);
      const uint32_t te = (uint32_t)(tb + (u_end - u) < ntile ? tb + (u_end - u) : ntile);
      split = bid_x() - (uint32_t)((uint64_t)qb * ntile / units_per_wg);
      nspl = cols_per_split;
      c0 = tb * TC;
      c1 = te * TC < n ? te * TC : n;
      len = te - tb;
      whole = te == ntile;  // the block's last piece
    } else {
      qb = bid_x();
      split = bid_y();
      nspl = gdim_y();
      c0 = split * cols_per_split;
      c1 = c0 + cols_per_split < n ? c0 + cols_per_split : n;
      len = 1;
    }
    if (!first) __syncthreads();  // every wave is done with the previous piece's tiles
    scan_block<PREC, J, KPR, NW>(tiles, xsq, xs32, dp32, n, kp, rowlist, nq, qb, tinit, split, nspl, c0, c1,
                                 eps_pre, cand_idx, cand_sc, cand_cnt, cand_floor, inserts);
    if (whole) {
      const uint32_t qq = qb * QB + tid_x();
      if (qq < nq)
        for (uint32_t s = split + 1; s < nspl; ++s) {
          cand_cnt[(size_t)qq * nspl + s] = 0;
          cand_floor[(size_t)qq * nspl + s] = -INFINITY;
        }
    }
    u += len;
  }
}

// items[3e..3e+2] = (sample, query block, split), in decreasing cost order
template <int PREC, int J, int KPR, int NW>
__global__ __launch_bounds__(64 * NW, (scan_waves<J, KPR>())) void coarse_batch_kernel(
    const BatchSample* __restrict__ samples, const uint32_t* __restrict__ items, uint32_t nitems,
    uint32_t* __restrict__ queue, const uint8_t* __restrict__ xsq, const float* __restrict__ xs32, uint32_t dp32,
    float eps_pre, uint32_t* __restrict__ cand_idx, float* __restrict__ cand_sc, uint32_t* __restrict__ cand_cnt,
    float* __restrict__ cand_floor, unsigned long long* __restrict__ inserts,
    const unsigned long long* __restrict__ bad) {
  __shared__ __attribute__((aligned(1024))) uint8_t tiles[2 * TileShape<J>::TILEB];
  __shared__ uint32_t s_item;
  // a sample with a NaN or inf row fails the whole call: no score is compared
  if (*bad != ~0ull) return;
  for (;;) {
    if (tid_x() == 0) s_item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it >= nitems) break;
    const uint32_t sidx = items[3 * it], qb = items[3 * it + 1], sp = items[3 * it + 2];
    const BatchSample& S = samples[sidx];
    const size_t rowb = (size_t)S.row_pad;
    const uint32_t c0 = sp * S.cols_per_split;
    const uint32_t c1 = c0 + S.cols_per_split < S.n ? c0 + S.cols_per_split : S.n;
    scan_block<PREC, J, KPR, NW>(tiles, xsq + rowb * 32 * J, xs32 + rowb * dp32, dp32, S.n, S.kp, nullptr, S.n,
                                 qb, nullptr, sp, S.nsplit, c0, c1, eps_pre, cand_idx + S.entry_base,
                                 cand_sc + S.entry_base, cand_cnt + S.slot_base, cand_floor + S.slot_base, inserts);
    __syncthreads();
  }
}

}  // namespace
