// scan.h -- the fused MFMA tile + candidate-list scan (knn.rs:603-660, 759-875), shared by the
// instantiation units (scan_inst.hip, one object per precision form and list capacity) and the
// host code that launches it.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "../../include/costal_hip.h"
#include "device_ids.h"
#include "order.h"

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

// The scan's inputs beyond the rows.  The pruned scan (order.h): the working order of the rows
// and, per query block, the candidate tiles in decreasing bound order; all null: candidates in row
// order, no pruning.  Tanimoto selection: the rows' raw norms and their ranges per half LDS tile.
struct ScanArgs {
  const uint32_t* perm;          // working row -> original row (null: the identity)
  const uint32_t* iperm;         // original row -> working row (for a row list)
  const uint32_t* seq;           // query block b: tiles seq[b*npt ..] in decreasing bound order
  const float* sb;               // their recorded bounds (non-increasing; each bounds every query-
                                 // candidate score of its tile and of the tiles after it)
  uint32_t npt;                  // candidate tiles
  uint32_t pt;                   // rows of a candidate tile (a multiple of the LDS tile)
  unsigned long long* scanned;   // candidate tiles scanned (summed over blocks and splits)
  const float* tnorm;            // Tanimoto: raw norm of every working row (f32; 0 past n)
  const float2* trange;          // Tanimoto: (smallest, largest) raw norm of every block of TC/2
                                 // working rows, rounded outward to f32
};

typedef void (*CoarseFn)(const uint8_t*, const float*, uint32_t, uint32_t, int, const uint32_t*,
                         uint32_t, const float*, uint32_t, uint32_t, float, uint32_t*, float*, uint32_t*,
                         float*, unsigned long long*, ScanArgs);
typedef void (*CoarseBatchFn)(const BatchSample*, const uint32_t*, uint32_t, uint32_t*, const uint8_t*,
                              const float*, uint32_t, float, uint32_t*, float*, uint32_t*, float*,
                              unsigned long long*, const unsigned long long*);

// One instantiation unit per (precision form, list capacity): the kernel for J 16-byte row chunks
// per lane half, or null when the form has none for that J.  Tanimoto selection has units of its
// own, split FP16 only (the coarse form never changes the output).
#define COSTAL_SCAN_DECL(P, K)                       \
  CoarseFn scan_fn_p##P##_k##K(uint32_t J);          \
  CoarseBatchFn scan_batch_fn_p##P##_k##K(uint32_t J);
COSTAL_SCAN_DECL(0, 16) COSTAL_SCAN_DECL(0, 32) COSTAL_SCAN_DECL(0, 64)
COSTAL_SCAN_DECL(1, 16) COSTAL_SCAN_DECL(1, 32) COSTAL_SCAN_DECL(1, 64)
COSTAL_SCAN_DECL(2, 16) COSTAL_SCAN_DECL(2, 32) COSTAL_SCAN_DECL(2, 64)
#undef COSTAL_SCAN_DECL
CoarseFn scan_tani_fn_k16(uint32_t J);
CoarseFn scan_tani_fn_k32(uint32_t J);
CoarseFn scan_tani_fn_k64(uint32_t J);

}  // namespace costal

namespace {

using costal::BatchSample;
using costal::ScanArgs;

constexpr int PREC_F32 = COSTAL_PREC_FP32;
constexpr int PREC_F16 = COSTAL_PREC_FP16;
constexpr uint32_t NONE = 0xffffffffu;
// selection metric of a scan instance
constexpr int SEL_COS = 0, SEL_TANI = 1;
// Tanimoto cut shave: the computed cosine cut c_tau * s_min is at most ~10 f32 roundings above the
// exact one; scaling it by 1 - 2^-18 puts it below.
constexpr float TANI_SHAVE = 1.0f - 0x1p-18f;

// Tanimoto selection on normalized rows.  With a, b the raw norms of the query and the candidate,
// rho = b / a and s = rho + 1/rho >= 2, the Tanimoto score of cosine c is T = c a b / (a^2 + b^2 -
// c a b) = c / (s - c) (weight.rs:28-37), increasing in c with slope s / (s - c)^2 <= 2 for c <= 1:
// a list score off by e in the cosine is off by at most 2e / (1 - e)^2 in T.  f32 evaluation from
// the list score v: rho from b and 1/a (v_rcp_f32, 1 ulp), s, then v / (s - v) by a reciprocal:
// at most ~16 roundings of a value <= 1 + 3e.  A zero query (1/a taken as 0) or a zero candidate
// gives s = inf and T = 0, as the reference scores zero rows.
__device__ __forceinline__ float tani32(float v, float b, float ia) {
  const float rho = b * ia;
  const float s = rho + __builtin_amdgcn_rcpf(rho);
  return v * __builtin_amdgcn_rcpf(s - v);
}

// The smallest s over candidates with raw norms in [r.x, r.y] for a query with 1/a = ia: s(rho) is
// smallest at rho = 1 and grows away from it (a zero query: inf).
__device__ __forceinline__ float tani_smin(float2 r, float ia) {
  const float lo = r.x * ia, hi = r.y * ia;
  if (hi < 1.0f) return hi + __builtin_amdgcn_rcpf(hi);
  if (lo > 1.0f) return lo + __builtin_amdgcn_rcpf(lo);
  return 2.0f;
}

// The cosine cut of a Tanimoto threshold tau > 0 is c_tau * smin: T > tau needs c > tau / (1 + tau)
// * s (T >= -1/3, so 1 + tau > 0), and s >= smin.  c_tau is shaved below the exact value; -inf
// (no cut) for tau <= 0 (T of a negative cosine rises with s).  smin is 2 or more, or inf.
__device__ __forceinline__ float tani_ctau(float tau) {
  if (!(tau > 0.0f)) return -INFINITY;
  return tau * __builtin_amdgcn_rcpf(1.0f + tau) * TANI_SHAVE;
}

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
// Workgroup = NW waves; each wave owns 32 queries, held in registers for the whole launch.  The
// workgroup streams its candidate range through two LDS buffers of TC-row tiles (32 KB each),
// staged global -> registers -> LDS ahead of the compute.  For every 32-candidate
// sub-tile a wave issues the MFMAs for D[cand][query] = sum_k cand[k] * query[k]: lane l owns query
// (l % 32) and the 16 candidates 8*(i/4) + 4*(l/32) + (i%4), i = 0..15 (the 32x32 accumulator
// layout).  The scan takes the max of a step's sub-tiles (max3 chains) and compares once with the
// lane's threshold; only waves with a hit take the insert path.  MFMAs run one step ahead of the
// scan (see scan_block).
//
// Candidate lists live in registers: each lane keeps its own sorted list of the best kp (score,
// candidate) pairs among the candidates it sees (its half of every sub-tile), so the two lanes of a
// query never wait on each other.  Both lists go to the rescoring, which treats them like two
// candidate-range splits: a candidate outside both lists scored at most the larger of the two
// floors.
//
// Row layout (global and LDS): a row is 2*J 16-byte chunks; chunk q feeds lane half h = q % 2,
// slot j = q / 2.  Queries and candidates use the same chunk -> (lane half, slot) map, so each
// MFMA contracts matching feature indices (any consistent permutation of k leaves the sum a sum of
// the same products).  In LDS the chunks of one (sub-tile, slot, half) are stored for the 32 lanes
// contiguously (1 KB per wave-wide load), so each ds_read_b128 is conflict free.


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

// Tanimoto instances that fit two waves per SIMD without scratch when a scan step holds one
// sub-tile: lists of 16 with J <= 6, lists of 32 with J = 2.
template <int J, int KPR>
constexpr bool tani_two_waves() {
  return (KPR == 16 && J <= 6) || (KPR == 32 && J == 2);
}

template <int J, bool ONE_SUB = false>
struct TileShape {
  static constexpr int TC = J <= 2 ? 512 : J <= 4 ? 256 : 128;  // candidates per tile: 32 KB tiles
  static constexpr int TILEB = TC * 32 * J;
  // Tanimoto: a tile's raw norms (TC floats) and its halves' norm ranges (two float2), in LDS
  static constexpr int TANIF = TC + 4;
  static constexpr int SUBT = TC / 32;                         // 32-candidate sub-tiles per tile
  // sub-tiles per scan step (ONE_SUB: one, which frees the registers of a second step's
  // accumulators and operands)
  static constexpr int GS = (!ONE_SUB && SUBT >= 4 && J <= 4) ? 2 : 1;
  static constexpr int NSTEP = SUBT / GS;                      // scan steps per tile
};

// One work item of the scan: query block `qblock` (QB queries) against candidates [c_begin, c_end)
// of one sample, whose lists go to split `split` of the queries' nsplit list pairs.  tile0/tile1
// are the workgroup's two LDS tile buffers.
//
// Software pipeline over scan steps (GS sub-tiles each, NSTEP per tile, G counting steps over the
// whole candidate range): step G issues the MFMAs of step G+1, reads the candidate chunks of step
// G+3 from LDS, and then scans the accumulators of step G, so the matrix core works on the next
// step while the wave compares and inserts, and the LDS reads complete behind two steps of work.
// Accumulators and chunk registers are double buffered by the parity of G (NSTEP is even, so the
// parity is the step's parity within its tile).  The next tile is stored to the other LDS buffer,
// and the workgroup meets at the barrier, at step NSTEP-3 of each tile, before the first read of
// that tile.
template <int PREC, int J, int KPR, int NW, int SEL>
__device__ __forceinline__ void scan_block(
    uint8_t* __restrict__ tile0, uint8_t* __restrict__ tile1,
    const uint8_t* __restrict__ xsq,   // n_pad rows of 32*J bytes (the working copy of this form)
    const float* __restrict__ xs32,    // f32 working copy (dp32 per row), for the FP16 re-score
    uint32_t dp32, uint32_t n, int kp,
    const uint32_t* __restrict__ rowlist, uint32_t nq,  // queries: rowlist[0..nq) or rows 0..nq-1
    uint32_t qblock, const float* __restrict__ tinit,   // per-row admission threshold (by row) or null
    uint32_t split, uint32_t nsplit,
    uint32_t c_begin, uint32_t c_end,  // candidate range; c_begin a multiple of the tile, c_end <= n
    float eps_pre,                     // FP16 form: bound on |FP16 score - f32 score|
    uint32_t* __restrict__ cand_idx, float* __restrict__ cand_sc, uint32_t* __restrict__ cand_cnt,
    float* __restrict__ cand_floor, unsigned long long* __restrict__ inserts,
    const ScanArgs& pr, float* __restrict__ s_thr, float* __restrict__ s_nrm) {
  constexpr int QB = 32 * NW;
  constexpr int TC = TileShape<J>::TC;
  constexpr int ROWB = 32 * J;
  constexpr int NT = 64 * NW;
  constexpr int CHUNKS = TC * 2 * J;        // 16-byte chunks per tile
  constexpr int CPT = (CHUNKS + NT - 1) / NT;
  constexpr bool ONE_SUB = SEL == SEL_TANI && tani_two_waves<J, KPR>();
  constexpr int GS = TileShape<J, ONE_SUB>::GS;
  constexpr int NSTEP = TileShape<J, ONE_SUB>::NSTEP;
  constexpr int SUBT = TileShape<J>::SUBT;
  static_assert(NSTEP >= 2 && NSTEP % 2 == 0, "pipeline steps per tile");
  static_assert(GS * NSTEP * 32 == TC, "tile shape");
  static_assert((SUBT / 2) % GS == 0, "the sub-tiles of a scan step lie in one half of the tile");
  static_assert(SEL == SEL_COS || PREC != PREC_F16, "Tanimoto scans have no re-scoring form");
  constexpr bool TANI = SEL == SEL_TANI;

  const int tid = tid_x();
  const int w = tid >> 6, l = tid & 63, h = l >> 5, l32 = l & 31;
  const uint32_t qq = qblock * QB + w * 32 + l32;
  const bool qvalid = qq < nq;
  // the query's original row (thresholds, output slot) and its row in the working copy (query
  // chunks, self exclusion: candidates are working rows)
  uint32_t qo = 0, qg = 0;
  if (qvalid) {
    if (rowlist) {
      qo = rowlist[qq];
      qg = pr.iperm ? pr.iperm[qo] : qo;
    } else {
      qg = qq;
      qo = pr.perm ? pr.perm[qq] : qq;
    }
  }

  // the wave's queries, in registers for the whole launch
  uint4 qreg[J];
  {
    const uint8_t* qrow = xsq + (size_t)qg * ROWB;
#pragma unroll
    for (int j = 0; j < J; ++j) qreg[j] = *reinterpret_cast<const uint4*>(qrow + (2 * j + h) * 16);
  }
  // Empty list entries hold the lane's admission threshold (-inf without one; +inf for padding
  // queries, which so never take the insert path): the list minimum ls[0] is then always the
  // lane's threshold, and real entries (score above it) are the ones with an index.
  float t0 = qvalid ? (tinit ? tinit[qo] : -INFINITY) : INFINITY;
  // a NaN threshold (caller data or a model with NaN parameters) admits everything: the test reads
  // the bits, since this code is built without NaN semantics
  if ((__float_as_uint(t0) & 0x7fffffffu) > 0x7f800000u) t0 = -INFINITY;
  float ls[KPR];
  uint32_t li[KPR];
#pragma unroll
  for (int e = 0; e < KPR; ++e) {
    ls[e] = e < kp ? t0 : INFINITY;
    li[e] = NONE;
  }
  float th = ls[0];
  float thp = -INFINITY;  // the partner list's threshold, refreshed once per tile
  uint32_t my_inserts = 0;
  // Tanimoto: list scores are f32 Tanimoto scores (tani32); a candidate whose cosine is at most the
  // cut of the lane's threshold for its half of the tile (c_tau times the half's smallest s) can
  // not enter, so the max3 test stays on the cosines.  1/a of the query (0 for a zero row), c_tau
  // of the lane's threshold, the halves' smallest s.
  float qia = 0.0f, ctau = -INFINITY, smin0 = 2.0f, smin1 = 2.0f;
  if constexpr (TANI) {
    if (qvalid) {
      const float qa = pr.tnorm[qg];
      qia = qa > 0.0f ? __builtin_amdgcn_rcpf(qa) : 0.0f;
    }
  }

  // Candidate tiles.  Without pruning: the range [c_begin, c_end) in row order, one LDS tile after
  // another, or (a row list in the working order) every nsplit-th LDS tile of all rows, so that a
  // row's near neighbors, which sit close together, spread over all of its lists.  Pruned: the
  // split's share of the query block's candidate tiles in decreasing bound order (positions
  // split, split + nsplit, ...); an LDS tile is a part of one candidate tile (R = pt / TC of them
  // per candidate tile) or two candidate tiles of half its rows (H = 2), and the scan stops before
  // the first LDS tile whose first candidate tile's bound is below every list's threshold.  Each
  // LDS tile is two halves of TC/2 rows with first rows (x, y); a second half past the split's
  // last position numbers its candidates from n, so that none is admitted (and loads the first
  // half's rows again).
  const uint32_t* seq_b = pr.seq ? pr.seq + (size_t)qblock * pr.npt : nullptr;
  const float* sb_b = pr.seq ? pr.sb + (size_t)qblock * pr.npt : nullptr;
  const bool interleave = !seq_b && pr.perm && nsplit > 1;
  int ntile, R = 1;
  bool two = false;
  uint32_t npos = 0;
  if (seq_b) {
    c_end = n;
    npos = pr.npt > split ? (pr.npt - split + nsplit - 1) / nsplit : 0;
    two = pr.pt < (uint32_t)TC;  // pt = TC / 2
    R = two ? 1 : (int)(pr.pt / (uint32_t)TC);
    ntile = two ? (int)((npos + 1) / 2) : (int)npos * R;
  } else if (interleave) {
    c_end = n;
    const uint32_t all = (n + TC - 1) / TC;
    ntile = all > split ? (int)((all - split + nsplit - 1) / nsplit) : 0;
  } else {
    ntile = c_begin < c_end ? (int)((c_end - c_begin + TC - 1) / TC) : 0;
  }
  auto tile_base = [&](int T) -> uint2 {
    uint32_t a;
    if (seq_b) {
      if (two) {
        const uint32_t p0 = 2 * (uint32_t)T;
        a = seq_b[split + p0 * nsplit] * pr.pt;
        return make_uint2(a, p0 + 1 < npos ? seq_b[split + (p0 + 1) * nsplit] * pr.pt : n);
      }
      a = seq_b[split + (uint32_t)(T / R) * nsplit] * pr.pt + (uint32_t)(T % R) * TC;
    } else {
      a = interleave ? (split + (uint32_t)T * nsplit) * TC : c_begin + (uint32_t)T * TC;
    }
    return make_uint2(a, a + TC / 2);
  };

  // Tiles go global -> registers -> LDS: the registers are loaded with tile t+2 while tile t is
  // scanned (right after the workgroup has stored tile t+1 and met) and stored during tile t+1, so
  // a load has a whole tile of steps to arrive.  One register set, and a loop body of one tile
  // whose only exit is at its end, swapping the two LDS buffers: the compiler's wait counts then
  // stay exact at the loop head (a second staging set, or an exit from the middle of a body
  // unrolled by two, made it wait there for loads still in flight).
  static_assert(CPT <= 8, "tile staging uses at most eight 16-byte registers per thread");
  // Tiles of 512 rows may be two candidate tiles of 256: the halves load from their own rows (every
  // thread's chunk C lies in one half, since NT divides half a tile); smaller tiles are always one
  // contiguous range (the second half's first row is the first half's plus TC/2).
  constexpr bool SPLIT_HALVES = TC == 512;
  static_assert(!SPLIT_HALVES || (CHUNKS / 2) % NT == 0, "halves of a tile in whole chunks per thread");
  uint4 ra0 = {}, ra1 = {}, ra2 = {}, ra3 = {}, ra4 = {}, ra5 = {}, ra6 = {}, ra7 = {};
  // Tanimoto staging: NPER raw norms of tile rows tid*NPER.., and (threads 0, 1) a half's range.
  // The norms live in three LDS slots, tile t in slot t % 3: unlike the rows, which every wave has
  // read into registers before the barrier of their tile, the norms are read by the insert path
  // until the end of the tile, after that barrier, so the slot of tile t is refilled (with tile
  // t + 3, at tile t + 2) only once every wave has met at the barrier of tile t + 1, after its last
  // read of tile t.
  constexpr int NPER = TC >= NT ? TC / NT : 1;
  constexpr int TANIF = TileShape<J>::TANIF;
  static_assert(NPER <= 2, "norm staging");
  float2 rn = {}, rr = {};
  int ns_cur = 0;  // the norm slot of tile t
  // Thread e fills LDS slot e (consecutive lanes, consecutive 16-byte slots: conflict free); slot
  // e = ((sub*J + j)*2 + half)*32 + lane32 holds chunk 2j + half of tile row sub*32 + lane32.
#define LOAD_ONE(R, C)                                                                            \
  if constexpr (C < CPT) {                                                                        \
    const int e_ = tid + C * NT;                                                                  \
    if (CHUNKS % NT == 0 || e_ < CHUNKS) {                                                        \
      const int lr_ = e_ & 31, hh_ = (e_ >> 5) & 1, j_ = (e_ >> 6) % J, sub_ = (e_ >> 6) / J;     \
      /* the half of the tile: chunk C of every thread lies in one half (SPLIT_HALVES) */          \
      constexpr int hs_ = SPLIT_HALVES && C * NT >= CHUNKS / 2 ? 1 : 0;                           \
      const uint32_t off_ = (uint32_t)(((sub_ - hs_ * (SUBT / 2)) * 32 + lr_) * ROWB + (2 * j_ + hh_) * 16); \
      R##C = *reinterpret_cast<const uint4*>((hs_ ? srcb_ : srca_) + off_);                      \
    }                                                                                             \
  }
#define LOAD_REGS(R, HB)                                                                          \
  {                                                                                               \
    const uint8_t* srca_ = xsq + (size_t)(HB).x * ROWB;                                            \
    const uint8_t* srcb_ = SPLIT_HALVES ? xsq + (size_t)((HB).y < n ? (HB).y : (HB).x) * ROWB : srca_; \
    LOAD_ONE(R, 0) LOAD_ONE(R, 1) LOAD_ONE(R, 2) LOAD_ONE(R, 3)                                   \
    LOAD_ONE(R, 4) LOAD_ONE(R, 5) LOAD_ONE(R, 6) LOAD_ONE(R, 7)                                   \
    if constexpr (TANI) LOAD_TANI(HB)                                                             \
  }
  // the second half's first row (a contiguous tile: the first half's plus TC/2; a second half past
  // the split's last position: the first half's rows again, never admitted)
#define HALF1(HB) (SPLIT_HALVES ? ((HB).y < n ? (HB).y : (HB).x) : (HB).x + TC / 2)
#define LOAD_TANI(HB)                                                                             \
  {                                                                                               \
    const int r0_ = tid * NPER;                                                                   \
    if (r0_ < TC) {                                                                               \
      const uint32_t row_ = r0_ < TC / 2 ? (HB).x + r0_ : HALF1(HB) + (r0_ - TC / 2);              \
      if constexpr (NPER == 2) rn = *reinterpret_cast<const float2*>(pr.tnorm + row_);             \
      else rn.x = pr.tnorm[row_];                                                                 \
    }                                                                                             \
    if (tid < 2) rr = pr.trange[(tid ? HALF1(HB) : (HB).x) / (TC / 2)];                            \
  }
#define STORE_ONE(R, C, BUF)                                                                      \
  if constexpr (C < CPT) {                                                                        \
    const int e_ = tid + C * NT;                                                                  \
    if (CHUNKS % NT == 0 || e_ < CHUNKS) *reinterpret_cast<uint4*>(&BUF[e_ * 16]) = R##C;         \
  }
#define STORE_REGS(R, BUF, NSLOT) \
  { STORE_ONE(R, 0, BUF) STORE_ONE(R, 1, BUF) STORE_ONE(R, 2, BUF) STORE_ONE(R, 3, BUF)         \
    STORE_ONE(R, 4, BUF) STORE_ONE(R, 5, BUF) STORE_ONE(R, 6, BUF) STORE_ONE(R, 7, BUF)         \
    if constexpr (TANI) {                                                                         \
      float* nb_ = s_nrm + (NSLOT) * TANIF;                                                       \
      if (tid * NPER < TC) {                                                                      \
        if constexpr (NPER == 2) *reinterpret_cast<float2*>(nb_ + tid * 2) = rn;                  \
        else nb_[tid] = rn.x;                                                                     \
      }                                                                                           \
      if (tid < 2) *reinterpret_cast<float2*>(nb_ + TC + 2 * tid) = rr;                           \
    } }
  // LDS writes of this wave done, then everyone meets; no wait on the global loads in flight
#define TILE_BARRIER()                                                                            \
  {                                                                                               \
    __builtin_amdgcn_s_waitcnt(0xc07f); /* lgkmcnt(0) */                                          \
    __builtin_amdgcn_s_barrier();                                                                 \
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");                                        \
  }
  // the lane's chunks of scan step S of the tile in BUF, into chunk buffer BS
#define READ_B(BS, BUF, S)                                                                        \
  _Pragma("unroll") for (int u_ = 0; u_ < GS; ++u_)                                               \
    _Pragma("unroll") for (int j_ = 0; j_ < J; ++j_)                                              \
      bop[BS][u_][j_] = *reinterpret_cast<const uint4*>(                                          \
          (BUF) + (((((S) * GS + u_) * J + j_) * 2 + h) * 32 + l32) * 16);
#define MFMA_STEP(SLOT) \
  _Pragma("unroll") for (int u_ = 0; u_ < GS; ++u_) acc[SLOT][u_] = tile_mfma<PREC, J>(bop[SLOT][u_], qreg);

  const float* xq32 = xs32 + (size_t)qg * dp32;
  f32x16 acc[2][GS];
  uint4 bop[2][GS][J];
  constexpr int RD = 3;  // chunks of step G+RD are read at step G
  static_assert(NSTEP >= RD, "steps per tile");
  int t = 0;
  // first candidates of the halves of tile t (in cur), of tile t+1 (in the staging registers)
  // and of tile t+2
  uint2 b_cur = make_uint2(0, 0), b_nx = b_cur, b_nn = b_cur;
  // Pruning, once per tile: every wave posts the lowest admission threshold of its lists (the
  // list minimum or the partner's, whichever is higher: what the list turns away) before the
  // barrier; after it, the workgroup's lowest threshold decides whether tile t+2 is scanned.
  // Thresholds only rise and the bounds fall along the order, so the first tile below it ends the
  // scan.  The slots alternate by tile, so a wave never overwrites a value another wave is still
  // to read.
  auto wave_threshold = [&]() __attribute__((always_inline)) {
    float m = fmaxf(th, thp);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fminf(m, __shfl_xor(m, off));
    if (l == 0) s_thr[(t & 1) * NW + w] = m;
  };
  auto prune_next = [&]() __attribute__((always_inline)) {
    float tb = s_thr[(t & 1) * NW];
#pragma unroll
    for (int v = 1; v < NW; ++v) tb = fminf(tb, s_thr[(t & 1) * NW + v]);
    const int T = t + 2;
    const uint32_t pos = two ? 2 * (uint32_t)T : (uint32_t)(T / R);
    if (T < ntile && (two || T % R == 0) && sb_b[split + pos * nsplit] < tb) ntile = T;
  };
  // compare the accumulators of step `p` of tile t with the lane's threshold; rare path: each lane
  // walks its own hits, sub-tile by sub-tile
  auto scan_step = [&](const f32x16 (&a)[GS], int p) __attribute__((always_inline)) {
    float mu[GS];
#pragma unroll
    for (int u = 0; u < GS; ++u) mu[u] = max16(a[u]);
    float m = mu[0];
#pragma unroll
    for (int u = 1; u < GS; ++u) m = fmaxf(m, mu[u]);
    const bool half1 = p * GS >= SUBT / 2;  // the step's half of the tile
    float the = TANI ? ctau * (half1 ? smin1 : smin0) : (PREC == PREC_F16) ? fmaxf(th, thp) - eps_pre : fmaxf(th, thp);
    if (__builtin_amdgcn_ballot_w64(m > the) != 0) {
#pragma unroll
      for (int u = 0; u < GS; ++u) {
        if (mu[u] > the) {
          // bit i = accumulator i above the threshold (built from the top, no wide constants)
          uint32_t mask = 0;
#pragma unroll
          for (int i = 15; i >= 0; --i) mask = (mask << 1) | (a[u][i] > the ? 1u : 0u);
          // the lane's offset is added here, not folded into per-sub-tile bases hoisted out of the
          // tile loop (each would hold a register for the whole launch)
          uint32_t lane_off = 4 * h;
          __asm__ volatile("" : "+v"(lane_off));
          const int sub = p * GS + u;  // the sub-tile, in the first or second half of the tile
          const uint32_t cbase =
              (sub < SUBT / 2 ? b_cur.x + (uint32_t)sub * 32 : b_cur.y + (uint32_t)(sub - SUBT / 2) * 32) + lane_off;
          while (mask) {
            const int i = __builtin_ctz(mask);
            mask &= mask - 1;
            float v = pick16(a[u], i);
            const uint32_t cand = cbase + 8 * (i >> 2) + (i & 3);
            if (v > the && cand < c_end && cand != qg) {
              if constexpr (PREC == PREC_F16) {
                v = dot32(xq32, xs32 + (size_t)cand * dp32, dp32);
                if (!(v > th)) continue;
              }
              if constexpr (TANI) {
                const float b = s_nrm[ns_cur * TANIF + sub * 32 + (int)lane_off + 8 * (i >> 2) + (i & 3)];
                v = tani32(v, b, qia);
                if (!(v > fmaxf(th, thp))) continue;
              }
              list_push<KPR>(ls, li, v, cand);
              th = ls[0];
              if constexpr (TANI) {
                ctau = tani_ctau(fmaxf(th, thp));
                the = ctau * (half1 ? smin1 : smin0);
              } else {
                the = (PREC == PREC_F16) ? fmaxf(th, thp) - eps_pre : fmaxf(th, thp);
              }
              ++my_inserts;
            }
          }
        }
      }
    }
  };
  // all NSTEP steps of tile t, held in CUR; the next tile goes from the staging registers to OTH
#define RUN_TILE(CUR, OTH)                                                                        \
  _Pragma("unroll") for (int p = 0; p < NSTEP; ++p) {                                             \
    MFMA_STEP((p + 1) & 1)                                                                        \
    if (p == 0) {                                                                                 \
      thp = partner32(th);                                                                        \
      if constexpr (TANI) {                                                                       \
        const f32x4 rg_ = *reinterpret_cast<const f32x4*>(s_nrm + ns_cur * TANIF + TC);           \
        smin0 = tani_smin(make_float2(rg_[0], rg_[1]), qia);                                      \
        smin1 = tani_smin(make_float2(rg_[2], rg_[3]), qia);                                      \
        ctau = tani_ctau(fmaxf(th, thp));                                                         \
      }                                                                                           \
    }                                                                                             \
    if (p == NSTEP - RD) {                                                                        \
      STORE_REGS(ra, OTH, ns_cur == 2 ? 0 : ns_cur + 1);                                          \
      if (seq_b) wave_threshold();                                                                \
      TILE_BARRIER();                                                                             \
      if (seq_b) prune_next();                                                                    \
      b_nn = tile_base(t + 2 < ntile ? t + 2 : ntile - 1);                                        \
      LOAD_REGS(ra, b_nn);                                                                        \
    }                                                                                             \
    if (p + RD < NSTEP) {                                                                         \
      READ_B((p + RD) & 1, CUR, p + RD)                                                           \
    } else {                                                                                      \
      READ_B((p + RD) & 1, OTH, p + RD - NSTEP)                                                   \
    }                                                                                             \
    scan_step(acc[p & 1], p);                                                                     \
  }

  if (ntile > 0) {
    uint8_t* cur = tile0;
    uint8_t* oth = tile1;
    b_cur = tile_base(0);
    LOAD_REGS(ra, b_cur);
    STORE_REGS(ra, cur, 0);
    TILE_BARRIER();
    b_nx = tile_base(ntile > 1 ? 1 : 0);
    LOAD_REGS(ra, b_nx);
    READ_B(0, cur, 0)
    READ_B(1, cur, 1)
    MFMA_STEP(0)
    READ_B(0, cur, 2)
    for (;;) {
      // tile t in cur, ra = tile t+1 in flight
      RUN_TILE(cur, oth)
      if (++t >= ntile) break;
      uint8_t* const tmp = cur;
      cur = oth;
      oth = tmp;
      b_cur = b_nx;
      b_nx = b_nn;
      ns_cur = ns_cur == 2 ? 0 : ns_cur + 1;
    }
  }
  if (seq_b && pr.scanned && tid == 0) {
    const uint32_t done = two ? (2 * (uint32_t)ntile < npos ? 2 * (uint32_t)ntile : npos) : (uint32_t)((ntile + R - 1) / R);
    atomicAdd(pr.scanned, (unsigned long long)done);
  }
#undef RUN_TILE
#undef MFMA_STEP
#undef READ_B
#undef TILE_BARRIER
#undef LOAD_REGS
#undef STORE_REGS
#undef LOAD_ONE
#undef STORE_ONE
#undef LOAD_TANI
#undef HALF1

  if (inserts) {
    // one atomic per wave
    uint32_t tot = my_inserts;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (l == 0) atomicAdd(inserts, (unsigned long long)tot);
  }
  const float thp_end = partner32(th);
  if (qvalid) {
    // lists by original row (a row list: by position in it), entries as original rows
    const size_t slot = ((size_t)(rowlist ? qq : qo) * nsplit + split) * 2 + h;
    const size_t base = slot * (size_t)kp;
    uint32_t c = 0;
#pragma unroll
    for (int e = 0; e < KPR; ++e) {
      if (e < kp && li[e] != NONE) {
        cand_idx[base + c] = pr.perm ? pr.perm[li[e]] : li[e];
        cand_sc[base + c] = ls[e];
        ++c;
      }
    }
    cand_cnt[slot] = c;
    cand_floor[slot] = fmaxf(th, thp_end);  // everything this list turned away scored at most this
  }
}

// Two launch shapes:
//   units_per_wg == 0: grid (query blocks, splits); workgroup (x, y) scans query block x against
//     columns [y * cols_per_split, (y+1) * cols_per_split).
//   units_per_wg > 0: one workgroup per resident slot, each taking the same number of (query block,
//     tile) units in query-block-major order, so every slot ends together.  A query block whose
//     units straddle two workgroups is scanned in two pieces: the one from tile 0 fills list pair 0,
//     the other list pair 1 (nsplit = 2); a block scanned whole marks pair 1 empty.
// Waves per SIMD the scan kernel is built for: two where the registers allow it without scratch
// (the compiler otherwise lets such instances grow past 256 registers by a few, which halves the
// occupancy: the split-FP16 headline instance measured 12.3 ms unbounded against 6.2 ms bounded);
// left to the compiler where two would spill: J of 3 or 4, lists of 64, and lists of 32 with J of
// 2, 7 or 8.  The Tanimoto scan holds a few more registers (1/a, the halves' cuts, norm staging):
// two waves where one sub-tile per scan step makes room (tani_two_waves), else one.
template <int J, int KPR, int SEL = SEL_COS>
constexpr int scan_waves() {
  if constexpr (SEL == SEL_TANI) return tani_two_waves<J, KPR>() ? 2 : 1;
  return (KPR >= 64 || J == 3 || J == 4 || (KPR >= 32 && (J == 2 || J >= 7))) ? 1 : 2;
}

template <int PREC, int J, int KPR, int NW, int SEL = SEL_COS>
__global__ __launch_bounds__(64 * NW, (scan_waves<J, KPR, SEL>())) void coarse_kernel(
    const uint8_t* __restrict__ xsq, const float* __restrict__ xs32, uint32_t dp32, uint32_t n, int kp,
    const uint32_t* __restrict__ rowlist, uint32_t nq, const float* __restrict__ tinit, uint32_t cols_per_split,
    uint32_t units_per_wg, float eps_pre, uint32_t* __restrict__ cand_idx, float* __restrict__ cand_sc,
    uint32_t* __restrict__ cand_cnt, float* __restrict__ cand_floor, unsigned long long* __restrict__ inserts,
    ScanArgs pr) {
  __shared__ __attribute__((aligned(16))) uint8_t tile0[TileShape<J>::TILEB];
  __shared__ __attribute__((aligned(16))) uint8_t tile1[TileShape<J>::TILEB];
  __shared__ __attribute__((aligned(16))) float s_nrm[SEL == SEL_TANI ? 3 * TileShape<J>::TANIF : 4];
  __shared__ float s_thr[2 * NW];
  constexpr uint32_t TC = TileShape<J>::TC;
  constexpr uint32_t QB = 32 * NW;
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
      split = tb == 0 ? 0 : 1;
      nspl = 2;
      c0 = tb * TC;
      c1 = te * TC < n ? te * TC : n;
      len = te - tb;
      whole = tb == 0 && te == ntile;
    } else {
      qb = bid_x();
      split = bid_y();
      nspl = gdim_y();
      c0 = split * cols_per_split;
      c1 = c0 + cols_per_split < n ? c0 + cols_per_split : n;
      len = 1;
    }
    if (!first) __syncthreads();  // every wave is done with the previous piece's tiles
    scan_block<PREC, J, KPR, NW, SEL>(tile0, tile1, xsq, xs32, dp32, n, kp, rowlist, nq, qb, tinit, split, nspl, c0, c1,
                                 eps_pre, cand_idx, cand_sc, cand_cnt, cand_floor, inserts, pr, s_thr, s_nrm);
    if (whole) {
      const uint32_t qq = qb * QB + (tid_x() >> 6) * 32 + (tid_x() & 31);
      if (qq < nq) {
        const size_t slot = ((size_t)qq * 2 + 1) * 2 + ((tid_x() >> 5) & 1);
        cand_cnt[slot] = 0;
        cand_floor[slot] = -INFINITY;
      }
    }
    u += len;
  }
}

// items[3e..3e+2] = (sample, query block, split), in decreasing cost order
template <int PREC, int J, int KPR, int NW>
__global__ __launch_bounds__(64 * NW) void coarse_batch_kernel(
    const BatchSample* __restrict__ samples, const uint32_t* __restrict__ items, uint32_t nitems,
    uint32_t* __restrict__ queue, const uint8_t* __restrict__ xsq, const float* __restrict__ xs32, uint32_t dp32,
    float eps_pre, uint32_t* __restrict__ cand_idx, float* __restrict__ cand_sc, uint32_t* __restrict__ cand_cnt,
    float* __restrict__ cand_floor, unsigned long long* __restrict__ inserts,
    const unsigned long long* __restrict__ bad) {
  __shared__ __attribute__((aligned(16))) uint8_t tile0[TileShape<J>::TILEB];
  __shared__ __attribute__((aligned(16))) uint8_t tile1[TileShape<J>::TILEB];
  __shared__ float s_thr[2 * NW];
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
    scan_block<PREC, J, KPR, NW, SEL_COS>(tile0, tile1, xsq + rowb * 32 * J, xs32 + rowb * dp32, dp32, S.n, S.kp, nullptr,
                                 S.n, qb, nullptr, sp, S.nsplit, c0, c1, eps_pre,
                                 cand_idx + S.entry_base, cand_sc + S.entry_base, cand_cnt + S.slot_base,
                                 cand_floor + S.slot_base, inserts, ScanArgs{}, s_thr, nullptr);
    __syncthreads();
  }
}

}  // namespace
