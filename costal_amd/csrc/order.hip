// order.hip -- locality order of the rows, cones of row groups and the per-block bound order that
// the pruned fused scan follows (see order.h for the argument).  Every kernel here is deterministic:
// the k-means sums are fixed-point integers (order-independent), ties go to the lowest index, and
// the sort is stable, so a call's working order depends on its input only.
#include "order.h"

#include <math.h>


#include <algorithm>
#include <type_traits>

#include "device_ids.h"

namespace costal {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr double KM_SCALE = 1099511627776.0;  // 2^40: fixed-point scale of the center sums
constexpr uint32_t NO_CENTER = 0xffffffffu;
// Angle added to every cone: covers the distance between a normalized f32 row and the exact unit
// vector of the original row (at most about 2^-24 rad) with a wide margin.
constexpr double CONE_MARGIN = 1e-6;
// Added to every bound: the f64 evaluation error of the bound and of the rescored cosine.
constexpr double BOUND_MARGIN = 1e-6;

__device__ __forceinline__ double wave_sum(double x) {
  // fixed butterfly: the same order on every run
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmin(x, __shfl_xor(x, off));
  return x;
}

// One row of dp floats into DPM registers (zero beyond dp).
template <int DPM>
__device__ __forceinline__ void load_row(const float* __restrict__ row, uint32_t dp, float (&v)[DPM]) {
#pragma unroll
  for (int f = 0; f < DPM; f += 4) {
    if ((uint32_t)f < dp) {
      const f32x4 q = *reinterpret_cast<const f32x4*>(row + f);
      v[f] = q[0], v[f + 1] = q[1], v[f + 2] = q[2], v[f + 3] = q[3];
    } else {
      v[f] = v[f + 1] = v[f + 2] = v[f + 3] = 0.0f;
    }
  }
}

// The center (LDS, stride DPM) of largest inner product with v; ties to the lowest index; NO_CENTER
// for a zero row.
template <int DPM>
__device__ __forceinline__ uint32_t nearest(const float (&v)[DPM], const float* __restrict__ s_cen, uint32_t C) {
  float nz = 0.0f;
#pragma unroll
  for (int f = 0; f < DPM; ++f) nz = fmaxf(nz, fabsf(v[f]));
  if (nz == 0.0f) return NO_CENTER;
  float best = -INFINITY;
  uint32_t bc = 0;
  for (uint32_t c = 0; c < C; ++c) {
    const float* cen = s_cen + c * DPM;
    float s = 0.0f;
#pragma unroll
    for (int f = 0; f < DPM; ++f) s = __builtin_fmaf(v[f], cen[f], s);
    if (s > best) {
      best = s;
      bc = c;
    }
  }
  return bc;
}

// centers[c] = sample row floor(c * sample / C)
__global__ void km_init_kernel(const float* __restrict__ xs32, uint32_t dp, uint32_t C, uint32_t sample,
                               uint32_t stride, float* __restrict__ centers) {
  const uint32_t c = bid_x(), f = tid_x();
  if (f >= dp) return;
  const uint64_t r = (uint64_t)c * sample / C * stride;
  centers[(size_t)c * dp + f] = xs32[r * dp + f];
}

// One k-means assignment pass over the sample: every sample row adds itself (fixed point) to the
// sum of its nearest center.
template <int DPM>
__global__ __launch_bounds__(256) void km_step_kernel(const float* __restrict__ xs32, uint32_t dp, uint32_t C,
                                                      uint32_t sample, uint32_t stride,
                                                      const float* __restrict__ centers, long long* __restrict__ sums,
                                                      uint32_t* __restrict__ counts) {
  extern __shared__ __attribute__((aligned(16))) uint8_t km_lds[];
  long long* s_sum = reinterpret_cast<long long*>(km_lds);
  float* s_cen = reinterpret_cast<float*>(s_sum + (size_t)C * DPM);
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_cen + (size_t)C * DPM);
  for (uint32_t e = tid_x(); e < C * DPM; e += 256) {
    const uint32_t c = e / DPM, f = e % DPM;
    s_sum[e] = 0;
    s_cen[e] = f < dp ? centers[(size_t)c * dp + f] : 0.0f;
  }
  for (uint32_t c = tid_x(); c < C; c += 256) s_cnt[c] = 0;
  __syncthreads();
  const uint32_t r = bid_x() * 256 + tid_x();
  if (r < sample) {
    float v[DPM];
    load_row<DPM>(xs32 + (size_t)r * stride * dp, dp, v);
    const uint32_t c = nearest<DPM>(v, s_cen, C);
    if (c != NO_CENTER) {
#pragma unroll
      for (int f = 0; f < DPM; ++f)
        if (v[f] != 0.0f)
          atomicAdd(reinterpret_cast<unsigned long long*>(&s_sum[c * DPM + f]),
                    (unsigned long long)(long long)rint((double)v[f] * KM_SCALE));
      atomicAdd(&s_cnt[c], 1u);
    }
  }
  __syncthreads();
  for (uint32_t e = tid_x(); e < C * DPM; e += 256) {
    const uint32_t c = e / DPM, f = e % DPM;
    if (f < dp && s_sum[e] != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(&sums[(size_t)c * dp + f]), (unsigned long long)s_sum[e]);
  }
  for (uint32_t c = tid_x(); c < C; c += 256)
    if (s_cnt[c]) atomicAdd(&counts[c], s_cnt[c]);
}

__global__ void km_clear_kernel(long long* __restrict__ sums, uint32_t* __restrict__ counts,
                                uint32_t* __restrict__ totals, uint32_t C, uint32_t dp) {
  const uint32_t e = bid_x() * bdim_x() + tid_x();
  if (e < C * dp) sums[e] = 0;
  if (e < C) counts[e] = 0;
  if (e <= C) totals[e] = 0;
}

// New centers: the normalized means of their rows (a center with no rows keeps its place); the
// sums and counts are cleared for the next pass.
// One wave per center, a lane per feature (dp <= 64): the new center is the normalized sum (kept
// when the center drew no row), and the sums and count are cleared for the next pass.  (A thread
// per center walking its features took 7-12 us per pass.)
__global__ __launch_bounds__(64) void km_update_kernel(uint32_t dp, long long* __restrict__ sums,
                                                       uint32_t* __restrict__ counts, float* __restrict__ centers) {
  const uint32_t c = bid_x(), f = tid_x();
  long long* s = sums + (size_t)c * dp;
  const double v = f < dp ? (double)s[f] : 0.0;
  const double nn = wave_sum(v * v);
  if (counts[c] > 0 && nn > 0.0 && f < dp) centers[(size_t)c * dp + f] = (float)(v * (1.0 / sqrt(nn)));
  if (f < dp) s[f] = 0;
  if (f == 0) counts[c] = 0;
}

// Every row's center (zero rows: center C, sorted last), its rank among the rows of its 256-row
// block with the same center, the per-block center histogram, stored center-major
// (hist[c * nb + block]), and the rows per center (totals, integer atomics).
template <int DPM>
__global__ __launch_bounds__(256) void km_assign_kernel(const float* __restrict__ xs32, uint32_t dp, uint32_t n,
                                                        uint32_t C, const float* __restrict__ centers,
                                                        uint32_t* __restrict__ assign, uint32_t* __restrict__ rank,
                                                        uint32_t* __restrict__ hist, uint32_t* __restrict__ totals) {
  extern __shared__ __attribute__((aligned(16))) uint8_t km_lds[];
  float* s_cen = reinterpret_cast<float*>(km_lds);
  uint32_t* s_h = reinterpret_cast<uint32_t*>(s_cen + (size_t)C * DPM);  // C + 1
  __shared__ uint32_t s_a[256];
  for (uint32_t e = tid_x(); e < C * DPM; e += 256) {
    const uint32_t c = e / DPM, f = e % DPM;
    s_cen[e] = f < dp ? centers[(size_t)c * dp + f] : 0.0f;
  }
  for (uint32_t c = tid_x(); c <= C; c += 256) s_h[c] = 0;
  __syncthreads();
  const uint32_t i = bid_x() * 256 + tid_x();
  uint32_t a = NO_CENTER;
  if (i < n) {
    float v[DPM];
    load_row<DPM>(xs32 + (size_t)i * dp, dp, v);
    a = nearest<DPM>(v, s_cen, C);
    if (a == NO_CENTER) a = C;
    atomicAdd(&s_h[a], 1u);
  }
  s_a[tid_x()] = a;
  __syncthreads();
  if (i < n) {
    uint32_t r = 0;
    for (uint32_t t = 0; t < tid_x(); ++t) r += s_a[t] == a ? 1u : 0u;
    assign[i] = a;
    rank[i] = r;
  }
  const uint32_t nb = gdim_x();
  for (uint32_t c = tid_x(); c <= C; c += 256) {
    hist[(size_t)c * nb + bid_x()] = s_h[c];
    if (s_h[c]) atomicAdd(&totals[c], s_h[c]);
  }
}

// First working row of every (center, block): the center's first row plus the exclusive prefix
// sum of its per-block counts (one workgroup per center), in place.
__global__ __launch_bounds__(256) void block_offsets_kernel(uint32_t* __restrict__ hist, uint32_t nb,
                                                            const uint32_t* __restrict__ start) {
  __shared__ uint32_t s[256];
  uint32_t* h = hist + (size_t)bid_x() * nb;
  const uint32_t t = tid_x();
  const uint32_t per = (nb + 255) / 256;
  const uint32_t lo = std::min(nb, t * per), hi = std::min(nb, lo + per);
  uint32_t sum = 0;
  for (uint32_t e = lo; e < hi; ++e) sum += h[e];
  s[t] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint32_t add = t >= off ? s[t - off] : 0u;
    __syncthreads();
    s[t] += add;
    __syncthreads();
  }
  uint32_t run = start[bid_x()] + s[t] - sum;
  for (uint32_t e = lo; e < hi; ++e) {
    const uint32_t x = h[e];
    h[e] = run;
    run += x;
  }
}

__global__ void km_scatter_kernel(const uint32_t* __restrict__ assign, const uint32_t* __restrict__ rank,
                                  const uint32_t* __restrict__ first, uint32_t n, uint32_t nb,
                                  uint32_t* __restrict__ perm, uint32_t* __restrict__ iperm) {
  const uint32_t i = bid_x() * bdim_x() + tid_x();
  if (i >= n) return;
  const uint32_t pos = first[(size_t)assign[i] * nb + i / 256] + rank[i];
  perm[pos] = i;
  iperm[i] = pos;
}

// Working copies in working order (hi = fp16(v), lo = fp16(v - hi), as the original-order prep).
__global__ void permute_kernel(const float* __restrict__ xs32, const uint32_t* __restrict__ perm, uint32_t n,
                               uint32_t n_pad, uint32_t dp, uint32_t dp16, int split16, float* __restrict__ xs32w,
                               _Float16* __restrict__ xs16w) {
  const uint32_t w = bid_x() * bdim_x() + tid_x();
  if (w >= n_pad) return;
  const float* src = w < n ? xs32 + (size_t)perm[w] * dp : nullptr;
  if (xs32w)
    for (uint32_t t = 0; t < dp; t += 4)
      *reinterpret_cast<f32x4*>(xs32w + (size_t)w * dp + t) = src ? *reinterpret_cast<const f32x4*>(src + t) : f32x4{};
  if (xs16w) {
    const uint32_t s16 = split16 ? 2 * dp16 : dp16;
    for (uint32_t t = 0; t < dp16; t += 8) {
      f16x8 hi, lo;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float v = (src && t + c < dp) ? src[t + c] : 0.0f;
        hi[c] = (_Float16)v;
        lo[c] = (_Float16)(v - (float)hi[c]);
      }
      *reinterpret_cast<f16x8*>(xs16w + (size_t)w * s16 + t) = hi;
      if (split16) *reinterpret_cast<f16x8*>(xs16w + (size_t)w * s16 + dp16 + t) = lo;
    }
  }
}

// Cone of positions [g*G, min(n, g*G + G)) of a row list (rows[w]: the original row at position w,
// e.g. the working order perm) (one wave): center = the normalized f64 sum of the rows, rounded to
// f32; angle = the largest angle between that f32 center and any row, in f64 with both vectors'
// own norms, plus CONE_MARGIN; pi when the sum is zero (the cone covers everything).  cen holds the
// center as an f64 unit vector (the f32 center over its f64 norm: the same direction), ang[3g ..]
// = (angle, its cosine, its sine).  With norms (Tanimoto), nr[2g ..] = the smallest and largest
// f64 row norm of the group.
template <int DPM>
__global__ __launch_bounds__(64) void cone_kernel(const float* __restrict__ xs32, uint32_t dp,
                                                  const uint32_t* __restrict__ rows, uint32_t n, uint32_t G,
                                                  double* __restrict__ cen, double* __restrict__ ang,
                                                  const double* __restrict__ norms, double* __restrict__ nr) {
  __shared__ float s_c[DPM];
  const uint32_t g = bid_x(), l = tid_x();
  const uint32_t w0 = g * G, w1 = std::min(n, w0 + G);
  double acc[DPM];
#pragma unroll
  for (int f = 0; f < DPM; ++f) acc[f] = 0.0;
  double nmin = INFINITY, nmax = 0.0;
  for (uint32_t w = w0 + l; w < w1; w += 64) {
    float v[DPM];
    load_row<DPM>(xs32 + (size_t)rows[w] * dp, dp, v);
#pragma unroll
    for (int f = 0; f < DPM; ++f) acc[f] += (double)v[f];
    if (norms) {
      const double nv = norms[rows[w]];
      nmin = fmin(nmin, nv);
      nmax = fmax(nmax, nv);
    }
  }
  if (norms) {
    nmin = wave_min(nmin);
    nmax = -wave_min(-nmax);
    if (l == 0) nr[2 * g] = nmin, nr[2 * g + 1] = nmax;
  }
  double nn = 0.0;
#pragma unroll
  for (int f = 0; f < DPM; ++f) {
    acc[f] = wave_sum(acc[f]);
    nn += acc[f] * acc[f];
  }
  const double nrm = sqrt(nn);
  if (!(nrm > 0.0)) {
    if (l < dp) cen[(size_t)g * dp + l] = 0.0;
    if (l == 0) ang[3 * g] = M_PI, ang[3 * g + 1] = -1.0, ang[3 * g + 2] = 0.0;
    return;
  }
  if (l < DPM) {
    double mine = 0.0;
#pragma unroll
    for (int f = 0; f < DPM; ++f) mine = (uint32_t)f == l ? acc[f] : mine;
    s_c[l] = (float)(mine / nrm);
  }
  __syncthreads();
  double cn = 0.0;
#pragma unroll
  for (int f = 0; f < DPM; ++f) cn += (double)s_c[f] * (double)s_c[f];
  cn = sqrt(cn);
  if (l < dp) cen[(size_t)g * dp + l] = (double)s_c[l] / cn;
  double m = 2.0;
  for (uint32_t w = w0 + l; w < w1; w += 64) {
    float v[DPM];
    load_row<DPM>(xs32 + (size_t)rows[w] * dp, dp, v);
    double dot = 0.0, vn = 0.0;
#pragma unroll
    for (int f = 0; f < DPM; ++f) {
      dot += (double)s_c[f] * (double)v[f];
      vn += (double)v[f] * (double)v[f];
    }
    // a zero row has cosine 0 with everything: it widens the cone to at least pi/2
    const double c = vn > 0.0 ? dot / (cn * sqrt(vn)) : 0.0;
    m = fmin(m, c);
  }
  m = wave_min(m);
  if (l == 0) {
    const double a = acos(fmax(-1.0, fmin(1.0, m))) + CONE_MARGIN;
    ang[3 * g] = a, ang[3 * g + 1] = cos(a), ang[3 * g + 2] = sin(a);
  }
}

// ------------------------------------------------------------------------------------------------
// Bitonic sort of one workgroup's values into descending order: 1024 threads, value s*1024 + t in
// thread t's register v[s] (s < E, E a power of two up to EMAX).  Exchange distances of 1024 and
// more stay inside a thread, those below 64 inside a wave (shuffles), and only the ones in between
// go through LDS (s_buf, 1024*E values), so a sort of 4096 values meets at 2*15 barriers.
template <class T>
__device__ __forceinline__ T shfl_xor_t(T x, int j) {
  if constexpr (sizeof(T) == 8) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    return (T)(((uint64_t)(uint32_t)__shfl_xor((int)hi, j) << 32) | (uint32_t)__shfl_xor((int)lo, j));
  } else {
    return (T)__shfl_xor((int)x, j);
  }
}

// the pair (a at index i, b at its partner i ^ j) of a pass of stage k: what index i keeps
template <class T>
__device__ __forceinline__ T keep(T a, T b, uint32_t i, uint32_t j, uint32_t k) {
  const bool lower = (i & j) == 0, desc = (i & k) == 0;
  const T hi = a > b ? a : b, lo = a > b ? b : a;
  return lower == desc ? hi : lo;
}

// distance SJ * 1024: the pairs (s, s | SJ) of one thread's registers
template <int SJ, int S = 0, int EMAX, class T>
__device__ __forceinline__ void reg_pass(T (&v)[EMAX], int E, uint32_t k, uint32_t t) {
  if constexpr (S < EMAX) {
    if constexpr ((S & SJ) == 0 && (S | SJ) < EMAX) {
      if ((S | SJ) < E) {
        const uint32_t i = (uint32_t)S * 1024 + t;
        const T a = v[S], b = v[S | SJ];
        v[S] = keep(a, b, i, (uint32_t)SJ * 1024, k);
        v[S | SJ] = keep(b, a, i + (uint32_t)SJ * 1024, (uint32_t)SJ * 1024, k);
      }
    }
    reg_pass<SJ, S + 1>(v, E, k, t);
  }
}

template <int EMAX, class T>
__device__ void block_sort_desc(T (&v)[EMAX], int E, T* s_buf) {
  const uint32_t t = tid_x();
  const uint32_t N = 1024u * (uint32_t)E;
  for (uint32_t k = 2; k <= N; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j >= 1024) {
        const uint32_t sj = j >> 10;
        if (sj == 1) reg_pass<1>(v, E, k, t);
        if constexpr (EMAX > 2) {
          if (sj == 2) reg_pass<2>(v, E, k, t);
        }
        if constexpr (EMAX > 4) {
          if (sj == 4) reg_pass<4>(v, E, k, t);
        }
      } else if (j >= 64) {
#pragma unroll
        for (int s = 0; s < EMAX; ++s)
          if (s < E) s_buf[s * 1024 + t] = v[s];
        __syncthreads();
#pragma unroll
        for (int s = 0; s < EMAX; ++s)
          if (s < E) {
            const uint32_t i = (uint32_t)s * 1024 + t;
            v[s] = keep(v[s], s_buf[i ^ j], i, j, k);
          }
        __syncthreads();
      } else {
#pragma unroll
        for (int s = 0; s < EMAX; ++s)
          if (s < E) {
            const uint32_t i = (uint32_t)s * 1024 + t;
            v[s] = keep(v[s], shfl_xor_t(v[s], (int)j), i, j, k);
          }
      }
    }
  }
}

// f32 as an unsigned key of the same order (non-NaN values), and back
__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// The largest Tanimoto score of a pair with cosine at most c and raw norms a in [amin, amax], b in
// [bmin, bmax]: T = c / (s - c) rises with c and, for c > 0, falls with s = rho + 1/rho (rho = b/a),
// which is smallest at the ratio nearest 1; a cosine <= 0 gives T <= 0.
__device__ __forceinline__ double tani_bound(double c, double amin, double amax, double bmin, double bmax) {
  if (!(c > 0.0) || amax == 0.0 || bmax == 0.0) return 0.0;
  const double lo = bmin / amax;
  const double hi = amin > 0.0 ? bmax / amin : INFINITY;
  double smin = 2.0;
  if (hi < 1.0) smin = hi + 1.0 / hi;
  else if (lo > 1.0) smin = lo + 1.0 / lo;
  const double cc = fmin(c, 1.0);
  return cc / (smin - cc);
}

// Tiles in decreasing bound order, up to buckets: every candidate tile's bound for query block
// bid_x() (an f32 rounded up) goes to one of NBUCKET buckets of the distance 1 + 2^-18 - bound, 16
// per octave (the f32 exponent and top four mantissa bits of that distance, which grow with it),
// and the tiles are placed by a counting pass, bucket by bucket from the highest bounds, in tile
// order within a bucket (ranks by wave ballots: deterministic).  Every tile's recorded bound is
// the largest bound in its bucket, so the recorded bounds never increase along the order and each
// is an upper bound of its tile's and every later tile's: the scan's stop at the first recorded
// bound below every list's threshold skips only tiles that cannot qualify (a few tiles of the
// bucket holding a threshold are scanned without need).  1024 threads; a sorting network over
// the same tiles took 0.25 ms on 396,460 rows.  With norm ranges (qnr, tnr), the bounds are on
// Tanimoto scores (tani_bound of the cosine bound).
constexpr uint32_t NBUCKET = 512;
constexpr uint32_t BUCKET_BITS = 9;

__device__ __forceinline__ uint32_t bound_bucket(float key) {
  const float u = (float)(1.0 + 0x1p-18 - (double)key);  // 2.6e-6 .. 2.0001 for bounds in [-1, 1 + 1e-6]
  const int q = (int)(__float_as_uint(u) >> 19) - (108 << 4);
  return (uint32_t)(q < 0 ? 0 : q >= (int)NBUCKET ? (int)NBUCKET - 1 : q);
}

__global__ __launch_bounds__(1024) void bound_sort_kernel(const double* __restrict__ qcen, const double* __restrict__ qang,
                                                          const double* __restrict__ tcen, const double* __restrict__ tang,
                                                          uint32_t dp, uint32_t npt, uint32_t* __restrict__ seq,
                                                          float* __restrict__ sb, const double* __restrict__ qnr,
                                                          const double* __restrict__ tnr) {
  constexpr int EMAX = PRUNE_MAX_TILES / 1024;
  __shared__ double s_q[64];
  __shared__ uint32_t s_cnt[NBUCKET], s_max[NBUCKET], s_run[NBUCKET];
  __shared__ uint32_t s_wc[16][NBUCKET];  // per slot: tiles of each bucket in each wave
  const uint32_t b = bid_x(), t0 = tid_x(), wv = t0 >> 6, lane = t0 & 63;
  if (t0 < 64) s_q[t0] = t0 < dp ? qcen[(size_t)b * dp + t0] : 0.0;
  for (uint32_t e = t0; e < NBUCKET; e += 1024) s_cnt[e] = 0, s_max[e] = 0, s_run[e] = 0;
  for (uint32_t e = t0; e < 16 * NBUCKET; e += 1024) (&s_wc[0][0])[e] = 0;
  __syncthreads();
  const double aq = qang[3 * b], cq = qang[3 * b + 1], sq = qang[3 * b + 2];
  const int E = (int)((npt + 1023) / 1024);
  uint32_t bk[EMAX];
  // the bucket of tile s * 1024 + t0 (a lambda applied to each slot by hand: a pragma-unrolled loop
  // around the feature loop is not unrolled by the compiler)
  auto place = [&](int s) __attribute__((always_inline)) {
    bk[s] = NBUCKET;  // no tile
    const uint32_t t = (uint32_t)s * 1024 + t0;
    if (s < E && t < npt) {
      const double* tc = tcen + (size_t)t * dp;
      double dot = 0.0;
      for (uint32_t f = 0; f < dp; ++f) dot += s_q[f] * tc[f];
      // cos(max(0, th - a)) with a = aq + at, from the cosines and sines of the angles: th > a iff
      // cos th < cos a (both in [0, pi]); then cos(th - a) = cos th cos a + sin th sin a (a cone of
      // angle pi, a zero sum, covers everything)
      const double at = tang[3 * t], ct = tang[3 * t + 1], st = tang[3 * t + 2];
      double bound = 1.0;
      if (aq + at < M_PI) {
        const double ca = cq * ct - sq * st, sa = sq * ct + cq * st;
        const double cth = fmax(-1.0, fmin(1.0, dot));
        if (cth < ca) bound = cth * ca + sqrt(fmax(0.0, 1.0 - cth * cth)) * sa;
      }
      if (qnr) bound = tani_bound(bound, qnr[2 * b], qnr[2 * b + 1], tnr[2 * t], tnr[2 * t + 1]);
      bound += BOUND_MARGIN;
      float kf = (float)bound;
      if ((double)kf < bound) kf = nextafterf(kf, INFINITY);
      bk[s] = bound_bucket(kf);
      atomicAdd(&s_cnt[bk[s]], 1u);
      atomicMax(&s_max[bk[s]], float_key(kf));
    }
  };
  place(0);
  if constexpr (EMAX > 1) place(1);
  if constexpr (EMAX > 2) place(2);
  if constexpr (EMAX > 3) place(3);
  static_assert(EMAX <= 4, "four tiles per thread at most");
  __syncthreads();
  // first place of every bucket: exclusive prefix sum of the counts (threads < NBUCKET, one each)
  {
    const uint32_t c = t0 < NBUCKET ? s_cnt[t0] : 0u;
    uint32_t x = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, off);
      if ((int)lane >= off) x += y;
    }
    if (lane == 63 && t0 < NBUCKET) s_wc[0][wv] = x;  // wave totals (s_wc is clear, reused here)
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wv && w < NBUCKET / 64; ++w) before += s_wc[0][w];
    __syncthreads();
    if (t0 < NBUCKET) s_run[t0] = before + x - c;
    if (t0 < NBUCKET / 64) s_wc[0][t0] = 0;
    __syncthreads();
  }
  // the tiles of each slot in turn: a tile's place = its bucket's next free place + the tiles of
  // its bucket in earlier waves of the slot + those in earlier lanes of its wave
  for (int s = 0; s < E; ++s) {
    uint32_t bkt = 0;
#pragma unroll
    for (int z = 0; z < EMAX; ++z)
      if (z == s) bkt = bk[z];
    const bool real = bkt < NBUCKET;
    uint64_t same = __builtin_amdgcn_ballot_w64(real);
#pragma unroll
    for (uint32_t bit = 0; bit < BUCKET_BITS; ++bit) {
      const bool on = (bkt >> bit) & 1;
      const uint64_t bal = __builtin_amdgcn_ballot_w64(on);
      same &= on ? bal : ~bal;
    }
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint32_t r_in = (uint32_t)__builtin_popcountll(same & below);
    if (real && r_in == 0) s_wc[wv][bkt] = (uint32_t)__builtin_popcountll(same);
    __syncthreads();
    if (real) {
      uint32_t pos = s_run[bkt] + r_in;
      for (uint32_t w = 0; w < wv; ++w) pos += s_wc[w][bkt];
      seq[(size_t)b * npt + pos] = (uint32_t)s * 1024 + t0;
      sb[(size_t)b * npt + pos] = key_float(s_max[bkt]);
    }
    __syncthreads();
    for (uint32_t e = t0; e < NBUCKET; e += 1024) {
      uint32_t add = 0;
      for (uint32_t w = 0; w < 16; ++w) add += s_wc[w][e], s_wc[w][e] = 0;
      s_run[e] += add;
    }
    __syncthreads();
  }
}

// rows[0..nr) (original rows) into working order: their working positions sorted (one workgroup),
// mapped back through perm.  A value is ~position, so that the descending sort is ascending.
__global__ __launch_bounds__(1024) void sort_rows_kernel(uint32_t* __restrict__ rows, uint32_t nr, uint32_t p2,
                                                         const uint32_t* __restrict__ perm,
                                                         const uint32_t* __restrict__ iperm) {
  constexpr int EMAX = RESCAN_PRUNE_ROWS / 1024;
  __shared__ uint32_t s_buf[RESCAN_PRUNE_ROWS];
  const uint32_t t0 = tid_x();
  const int E = (int)(p2 > 1024 ? p2 / 1024 : 1);
  uint32_t v[EMAX];
#pragma unroll
  for (int s = 0; s < EMAX; ++s) {
    const uint32_t i = (uint32_t)s * 1024 + t0;
    v[s] = (s < E && i < nr) ? ~iperm[rows[i]] : 0u;  // padding (0) sorts last
  }
  block_sort_desc<EMAX>(v, E, s_buf);
#pragma unroll
  for (int s = 0; s < EMAX; ++s) {
    const uint32_t i = (uint32_t)s * 1024 + t0;
    if (s < E && i < nr) rows[i] = perm[~v[s]];
  }
}

// ------------------------------------------------------------------------------------------------
// Second level of the working order.

// Sub-centers of every first-level center: about one per `rows` of its rows, at least one, at most
// cap (the zero rows' bucket C: one), and their offsets (one workgroup, C + 1 <= 1024).
__global__ __launch_bounds__(1024) void sub_plan_kernel(const uint32_t* __restrict__ starts, uint32_t C, uint32_t n,
                                                        uint32_t rows, uint32_t cap, uint32_t* __restrict__ sub_cnt,
                                                        uint32_t* __restrict__ sub_off) {
  __shared__ uint32_t s[1024];
  const uint32_t c = tid_x();
  uint32_t m = 0;
  if (c <= C) {
    const uint32_t size = (c < C ? starts[c + 1] : n) - starts[c];
    m = c < C ? std::max<uint32_t>(1, std::min<uint32_t>(cap, (size + rows / 2) / rows)) : 1;
    sub_cnt[c] = m;
  }
  s[c] = m;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t add = c >= off ? s[c - off] : 0u;
    __syncthreads();
    s[c] += add;
    __syncthreads();
  }
  if (c <= C) sub_off[c] = s[c] - m;
}

// Spherical k-means of the sub-centers of first-level center bid_x() on a sample of its rows
// (working rows [starts[c], starts[c+1]) at a stride), deterministic as the first level.
template <int DPM>
__global__ __launch_bounds__(256) void sub_fit_kernel(const float* __restrict__ xs32, uint32_t dp,
                                                      const uint32_t* __restrict__ perm,
                                                      const uint32_t* __restrict__ starts, uint32_t C,
                                                      const uint32_t* __restrict__ sub_cnt,
                                                      const uint32_t* __restrict__ sub_off, uint32_t sample_max,
                                                      uint32_t iters, float* __restrict__ sub_cen) {
  extern __shared__ __attribute__((aligned(16))) uint8_t km_lds[];
  const uint32_t c = bid_x();
  const uint32_t m = sub_cnt[c];
  if (m <= 1) return;
  const uint32_t s0 = starts[c], size = starts[c + 1] - s0;
  const uint32_t S = std::min(size, sample_max), stride = std::max<uint32_t>(1, size / S);
  long long* s_sum = reinterpret_cast<long long*>(km_lds);
  float* s_cen = reinterpret_cast<float*>(s_sum + (size_t)m * DPM);
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_cen + (size_t)m * DPM);
  for (uint32_t e = tid_x(); e < m * DPM; e += 256) {
    const uint32_t j = e / DPM, f = e % DPM;
    const uint32_t row = perm[s0 + (uint64_t)j * S / m * stride];
    s_cen[e] = f < dp ? xs32[(size_t)row * dp + f] : 0.0f;
  }
  for (uint32_t it = 0; it < iters; ++it) {
    for (uint32_t e = tid_x(); e < m * DPM; e += 256) s_sum[e] = 0;
    for (uint32_t j = tid_x(); j < m; j += 256) s_cnt[j] = 0;
    __syncthreads();
    for (uint32_t r = tid_x(); r < S; r += 256) {
      float v[DPM];
      load_row<DPM>(xs32 + (size_t)perm[s0 + r * stride] * dp, dp, v);
      const uint32_t j = nearest<DPM>(v, s_cen, m);
      if (j != NO_CENTER) {
#pragma unroll
        for (int f = 0; f < DPM; ++f)
          if (v[f] != 0.0f)
            atomicAdd(reinterpret_cast<unsigned long long*>(&s_sum[j * DPM + f]),
                      (unsigned long long)(long long)rint((double)v[f] * KM_SCALE));
        atomicAdd(&s_cnt[j], 1u);
      }
    }
    __syncthreads();
    for (uint32_t j = tid_x(); j < m; j += 256) {
      if (s_cnt[j] == 0) continue;
      double nn = 0.0;
      for (uint32_t f = 0; f < dp; ++f) nn += (double)s_sum[j * DPM + f] * (double)s_sum[j * DPM + f];
      if (nn > 0.0) {
        const double inv = 1.0 / sqrt(nn);
        for (uint32_t f = 0; f < dp; ++f) s_cen[j * DPM + f] = (float)((double)s_sum[j * DPM + f] * inv);
      }
    }
    __syncthreads();
  }
  for (uint32_t e = tid_x(); e < m * DPM; e += 256) {
    const uint32_t j = e / DPM, f = e % DPM;
    if (f < dp) sub_cen[((size_t)sub_off[c] + j) * dp + f] = s_cen[e];
  }
}

// key2[w] = the global sub-center id of working row w (its first-level center's first sub-center
// plus its nearest sub-center there; ties to the lowest), shifted left by nbits and, with norms,
// or-ed with the row's norm key: floor((log2(norm) - l0) * lscale) clamped to [0, 2^nbits) (0 for
// a zero row).
template <int DPM>
__global__ __launch_bounds__(256) void sub_assign_kernel(const float* __restrict__ xs32, uint32_t dp,
                                                         const uint32_t* __restrict__ perm,
                                                         const uint32_t* __restrict__ assign, uint32_t n, uint32_t C,
                                                         const uint32_t* __restrict__ sub_cnt,
                                                         const uint32_t* __restrict__ sub_off,
                                                         const float* __restrict__ sub_cen, const double* __restrict__ norms,
                                                         double l0, double lscale, uint32_t nbits,
                                                         uint32_t* __restrict__ key2) {
  const uint32_t w = bid_x() * 256 + tid_x();
  if (w >= n) return;
  const uint32_t row = perm[w];
  const uint32_t c = assign[row];
  const uint32_t m = sub_cnt[c], base = sub_off[c];
  uint32_t g = base;
  if (c < C && m > 1) {
    float v[DPM];
    load_row<DPM>(xs32 + (size_t)row * dp, dp, v);
    float best = -INFINITY;
    for (uint32_t j = 0; j < m; ++j) {
      float u[DPM];
      load_row<DPM>(sub_cen + ((size_t)base + j) * dp, dp, u);
      float sc = 0.0f;
#pragma unroll
      for (int f = 0; f < DPM; ++f) sc = __builtin_fmaf(v[f], u[f], sc);
      if (sc > best) {
        best = sc;
        g = base + j;
      }
    }
  }
  uint32_t nk = 0;
  if (norms) {
    const double v = norms[row];
    if (v > 0.0) {
      const double q = floor((log2(v) - l0) * lscale);
      nk = q <= 0.0 ? 0u : q >= (double)((1u << nbits) - 1) ? (1u << nbits) - 1 : (uint32_t)q;
    }
  }
  key2[w] = (g << nbits) | nk;
}

// ------------------------------------------------------------------------------------------------
// Stable LSD radix sort of (key, value) pairs, RADIX_BITS per pass, by counting: per 256-row block
// the digit histogram (digit-major, hist[digit * nb + block]) and every row's rank among the
// block's rows of the same digit (wave ballots: lanes of equal digit, then the earlier waves'
// counts), the digits' first slots (exclusive scan of the totals), the blocks' first slots within
// each digit (block_offsets_kernel), then the scatter.
constexpr uint32_t RADIX_BITS = 8, RADIX_B = 1u << RADIX_BITS;

__global__ __launch_bounds__(256) void digit_count_kernel(const uint32_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                          uint32_t* __restrict__ rank, uint32_t* __restrict__ hist,
                                                          uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_cnt[4][RADIX_B];
  const uint32_t t = tid_x(), wv = t >> 6, lane = t & 63;
  for (uint32_t e = t; e < 4 * RADIX_B; e += 256) (&s_cnt[0][0])[e] = 0;
  __syncthreads();
  const uint32_t i = bid_x() * 256 + t;
  const bool valid = i < n;
  const uint32_t dg = valid ? (keys[i] >> shift) & (RADIX_B - 1) : 0;
  // lanes of this wave with the same digit (valid rows only)
  uint64_t same = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
  for (uint32_t b = 0; b < RADIX_BITS; ++b) {
    const bool bit = (dg >> b) & 1;
    const uint64_t bal = __builtin_amdgcn_ballot_w64(bit);
    same &= bit ? bal : ~bal;
  }
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t r_in = (uint32_t)__builtin_popcountll(same & below);
  if (valid && r_in == 0) s_cnt[wv][dg] = (uint32_t)__builtin_popcountll(same);
  __syncthreads();
  if (valid) {
    uint32_t r = r_in;
    for (uint32_t v = 0; v < wv; ++v) r += s_cnt[v][dg];
    rank[i] = r;
  }
  const uint32_t nb = gdim_x();
  for (uint32_t d = t; d < RADIX_B; d += 256) {
    const uint32_t c = s_cnt[0][d] + s_cnt[1][d] + s_cnt[2][d] + s_cnt[3][d];
    hist[(size_t)d * nb + bid_x()] = c;
    if (c) atomicAdd(&totals[d], c);
  }
}

__global__ void digit_scatter_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n,
                                     uint32_t shift, const uint32_t* __restrict__ rank,
                                     const uint32_t* __restrict__ first, uint32_t* __restrict__ keys_out,
                                     uint32_t* __restrict__ vals_out) {
  const uint32_t i = bid_x() * bdim_x() + tid_x();
  if (i >= n) return;
  const uint32_t k = keys[i];
  const uint32_t pos = first[(size_t)((k >> shift) & (RADIX_B - 1)) * gdim_x() + i / 256] + rank[i];
  keys_out[pos] = k;
  vals_out[pos] = vals[i];
}

__global__ void clear_kernel(uint32_t* __restrict__ v, uint32_t m) {
  const uint32_t e = bid_x() * bdim_x() + tid_x();
  if (e < m) v[e] = 0;
}

// In-place exclusive prefix sum of m <= 1024 counts (one workgroup).
__global__ __launch_bounds__(1024) void small_scan_kernel(uint32_t* __restrict__ v, uint32_t m) {
  __shared__ uint32_t s[1024];
  const uint32_t t = tid_x();
  const uint32_t x = t < m ? v[t] : 0u;
  s[t] = x;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t add = t >= off ? s[t - off] : 0u;
    __syncthreads();
    s[t] += add;
    __syncthreads();
  }
  if (t < m) v[t] = s[t] - x;
}

__global__ void invert_kernel(const uint32_t* __restrict__ perm, uint32_t n, uint32_t* __restrict__ iperm) {
  const uint32_t w = bid_x() * bdim_x() + tid_x();
  if (w < n) iperm[perm[w]] = w;
}

template <class F>
hipError_t by_dpm(uint32_t dp, F f) {
  if (dp <= 16) return f(std::integral_constant<int, 16>{});
  if (dp <= 32) return f(std::integral_constant<int, 32>{});
  if (dp <= 48) return f(std::integral_constant<int, 48>{});
  if (dp <= 64) return f(std::integral_constant<int, 64>{});
  return hipErrorInvalidValue;
}

inline uint32_t dpm_of(uint32_t dp) { return dp <= 16 ? 16 : dp <= 32 ? 32 : dp <= 48 ? 48 : 64; }

// rows per sub-center: one candidate tile (cosine), TANI_SUB_TILES of them (Tanimoto)
inline uint32_t sub_rows(const OrderShape& o) { return o.tani ? o.pt * TANI_SUB_TILES : o.pt; }

}  // namespace

OrderShape order_shape(uint32_t n, uint32_t dp, uint32_t qb, uint32_t clusters, bool tani, uint32_t tile) {
  OrderShape o{};
  o.tani = tani ? 1 : 0;
  o.norm_bits = tani ? TANI_NORM_BITS : 0;
  o.n = n;
  o.dp = dp;
  uint32_t C = clusters;
  if (!C) {
    // about 3,000 rows per center, a power of two in [16, 256]
    C = 16;
    while (C < 256 && (uint64_t)C * 2 * 3072 <= n) C *= 2;
  }
  const uint32_t cmax = KM_LDS_MAX / (KM_LDS_BYTES * dpm_of(dp)) - 1;
  o.C = std::max<uint32_t>(1, std::min<uint32_t>({C, cmax, std::max<uint32_t>(1, n / 4)}));
  o.sample = std::min<uint32_t>(n, 32768);
  o.iters = 5;
  // second level: one sub-center per candidate tile of rows (set below), fitted on at most 2048 rows
  o.sub_cap = std::min<uint32_t>(1024, KM_LDS_MAX / (KM_LDS_BYTES * dpm_of(dp)) - 1);
  // (the fit's effort barely matters: 0, 1 or 4 iterations, 512 or 2048 rows scan within 16.1-16.3%
  // of the tile pairs on 396,460 x 15)
  o.sub_sample = 512;
  o.sub_iters = 1;
  o.qb = qb;
  o.nblk = (n + qb - 1) / qb;
  o.pt = tile == 2 * PRUNE_TILE || (n + PRUNE_TILE - 1) / PRUNE_TILE > PRUNE_MAX_TILES ? 2 * PRUNE_TILE : PRUNE_TILE;
  o.npt = (n + o.pt - 1) / o.pt;
  // sub-centers: at most round(size / rows) per center (<= size / rows + 1/2), summed over C
  // centers, plus one for the zero rows
  o.G2 = n / sub_rows(o) + o.C + 2;
  o.key_bits = 1;
  while ((1ull << o.key_bits) < o.G2) ++o.key_bits;
  o.key_bits += o.norm_bits;
  return o;
}

bool order_supported(const OrderShape& o) {
  return o.dp <= 64 && o.npt >= 8 && o.npt <= PRUNE_MAX_TILES && o.key_bits <= 3 * 8;
}

namespace {
struct Carver {
  uint8_t* base;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
};

OrderBufs carve(Carver& c, const OrderShape& o) {
  const size_t nb = (o.n + 255) / 256;
  OrderBufs b{};
  b.perm = c.take<uint32_t>(o.n);
  b.iperm = c.take<uint32_t>(o.n);
  b.centers = c.take<float>((size_t)o.C * o.dp);
  b.sums = c.take<long long>((size_t)o.C * o.dp);
  b.counts = c.take<uint32_t>(o.C);
  b.assign = c.take<uint32_t>(o.n);
  b.rank = c.take<uint32_t>(o.n);
  b.hist = c.take<uint32_t>((size_t)std::max<uint32_t>(o.C + 1, RADIX_B) * nb);
  b.totals = c.take<uint32_t>(std::max<uint32_t>(o.C + 1, RADIX_B));
  b.sub_cnt = c.take<uint32_t>(o.C + 1);
  b.sub_off = c.take<uint32_t>(o.C + 1);
  b.sub_cen = c.take<float>((size_t)o.G2 * o.dp);
  b.key2 = c.take<uint32_t>(o.n);
  b.key2s = c.take<uint32_t>(o.n);
  b.perm2 = c.take<uint32_t>(o.n);
  b.rank2 = c.take<uint32_t>(o.n);
  b.qcen = c.take<double>((size_t)o.nblk * o.dp);
  b.qang = c.take<double>(3 * (size_t)o.nblk);
  b.tcen = c.take<double>((size_t)o.npt * o.dp);
  b.tang = c.take<double>(3 * (size_t)o.npt);
  b.qnr = c.take<double>(2 * (size_t)o.nblk);
  b.tnr = c.take<double>(2 * (size_t)o.npt);
  b.seq = c.take<uint32_t>((size_t)o.nblk * o.npt);
  b.sb = c.take<float>((size_t)o.nblk * o.npt);
  const size_t rb = (RESCAN_PRUNE_ROWS + o.qb - 1) / o.qb;
  b.rcen = c.take<double>(rb * o.dp);
  b.rang = c.take<double>(3 * rb);
  b.rnr = c.take<double>(2 * rb);
  b.rseq = c.take<uint32_t>(rb * o.npt);
  b.rsb = c.take<float>(rb * o.npt);
  return b;
}
}  // namespace

size_t order_bytes(const OrderShape& o) {
  Carver c{nullptr};
  carve(c, o);
  return c.off + 256;
}

OrderBufs order_carve(void* base, const OrderShape& o) {
  Carver c{reinterpret_cast<uint8_t*>(base)};
  return carve(c, o);
}

hipError_t order_rows(hipStream_t s, const float* xs32, const OrderShape& o, OrderBufs& b, const double* norms,
                      double nlo, double nhi) {
  const uint32_t stride = std::max<uint32_t>(1, o.n / o.sample);
  const uint32_t nb = (o.n + 255) / 256;
  return by_dpm(o.dp, [&](auto dpm) -> hipError_t {
    constexpr int DPM = decltype(dpm)::value;
    km_init_kernel<<<o.C, 64, 0, s>>>(xs32, o.dp, o.C, o.sample, stride, b.centers);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the sums start at zero (a kernel, not a memset: callers may capture the stream); the update
    // kernel clears them after every pass
    km_clear_kernel<<<(o.C * o.dp + 255) / 256, 256, 0, s>>>(b.sums, b.counts, b.totals, o.C, o.dp);
    const size_t lds_step = (size_t)o.C * DPM * 12 + o.C * 4;
    for (uint32_t it = 0; it < o.iters; ++it) {
      km_step_kernel<DPM><<<(o.sample + 255) / 256, 256, lds_step, s>>>(xs32, o.dp, o.C, o.sample, stride, b.centers,
                                                                          b.sums, b.counts);
      km_update_kernel<<<o.C, 64, 0, s>>>(o.dp, b.sums, b.counts, b.centers);
    }
    const size_t lds_assign = (size_t)o.C * DPM * 4 + (o.C + 1) * 4;
    km_assign_kernel<DPM><<<nb, 256, lds_assign, s>>>(xs32, o.dp, o.n, o.C, b.centers, b.assign, b.rank, b.hist,
                                                       b.totals);
    small_scan_kernel<<<1, 1024, 0, s>>>(b.totals, o.C + 1);
    block_offsets_kernel<<<o.C + 1, 256, 0, s>>>(b.hist, nb, b.totals);
    km_scatter_kernel<<<nb, 256, 0, s>>>(b.assign, b.rank, b.hist, o.n, nb, b.perm, b.iperm);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // second level: after the scatter, totals holds every center's first working row
    sub_plan_kernel<<<1, 1024, 0, s>>>(b.totals, o.C, o.n, sub_rows(o), o.sub_cap, b.sub_cnt, b.sub_off);
    const size_t lds_sub = (size_t)o.sub_cap * DPM * 12 + o.sub_cap * 4;
    sub_fit_kernel<DPM><<<o.C, 256, lds_sub, s>>>(xs32, o.dp, b.perm, b.totals, o.C, b.sub_cnt, b.sub_off,
                                                  o.sub_sample, o.sub_iters, b.sub_cen);
    // Tanimoto: norm keys log-spaced over [nlo, nhi] (the smallest positive and the largest norm)
    const bool nk = o.tani && norms;
    const double l0 = nk && nlo > 0.0 ? log2(nlo) : 0.0;
    const double lscale = nk && nhi > nlo && nlo > 0.0 ? (double)(1u << o.norm_bits) / (log2(nhi) - l0) : 0.0;
    sub_assign_kernel<DPM><<<nb, 256, 0, s>>>(xs32, o.dp, b.perm, b.assign, o.n, o.C, b.sub_cnt, b.sub_off, b.sub_cen,
                                              nk ? norms : nullptr, l0, lscale, nk ? o.norm_bits : 0u, b.key2);
    // stable: rows of one key keep the first level's order.  Passes of RADIX_BITS over the key's
    // bits (at least two), (key2, perm) -> (key2s, perm2) -> back ...; an odd count ends in perm2.
    const uint32_t passes = std::max<uint32_t>(2, (o.key_bits + RADIX_BITS - 1) / RADIX_BITS);
    uint32_t* kin = b.key2;
    uint32_t* vin = b.perm;
    uint32_t* kout = b.key2s;
    uint32_t* vout = b.perm2;
    for (uint32_t shift = 0; shift < passes * RADIX_BITS; shift += RADIX_BITS) {
      clear_kernel<<<1, RADIX_B, 0, s>>>(b.totals, RADIX_B);
      digit_count_kernel<<<nb, 256, 0, s>>>(kin, o.n, shift, b.rank2, b.hist, b.totals);
      small_scan_kernel<<<1, 1024, 0, s>>>(b.totals, RADIX_B);
      block_offsets_kernel<<<RADIX_B, 256, 0, s>>>(b.hist, nb, b.totals);
      digit_scatter_kernel<<<nb, 256, 0, s>>>(kin, vin, o.n, shift, b.rank2, b.hist, kout, vout);
      std::swap(kin, kout);
      std::swap(vin, vout);
    }
    if (vin != b.perm) {
      e = hipMemcpyAsync(b.perm, vin, (size_t)o.n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
    }
    invert_kernel<<<nb, 256, 0, s>>>(b.perm, o.n, b.iperm);
    return hipGetLastError();
  });
}

hipError_t order_permute(hipStream_t s, const float* xs32, const OrderShape& o, const OrderBufs& b, uint32_t n_pad,
                         uint32_t dp16, int split16, float* xs32w, _Float16* xs16w) {
  permute_kernel<<<(n_pad + 255) / 256, 256, 0, s>>>(xs32, b.perm, o.n, n_pad, o.dp, dp16, split16, xs32w, xs16w);
  return hipGetLastError();
}

hipError_t order_bounds(hipStream_t s, const float* xs32, const OrderShape& o, OrderBufs& b, const double* norms) {
  const double* nv = o.tani ? norms : nullptr;
  return by_dpm(o.dp, [&](auto dpm) -> hipError_t {
    constexpr int DPM = decltype(dpm)::value;
    cone_kernel<DPM><<<o.nblk, 64, 0, s>>>(xs32, o.dp, b.perm, o.n, o.qb, b.qcen, b.qang, nv, b.qnr);
    cone_kernel<DPM><<<o.npt, 64, 0, s>>>(xs32, o.dp, b.perm, o.n, o.pt, b.tcen, b.tang, nv, b.tnr);
    bound_sort_kernel<<<o.nblk, 1024, 0, s>>>(b.qcen, b.qang, b.tcen, b.tang, o.dp, o.npt, b.seq, b.sb,
                                              nv ? b.qnr : nullptr, nv ? b.tnr : nullptr);
    return hipGetLastError();
  });
}

hipError_t order_rescan(hipStream_t s, const float* xs32, const OrderShape& o, OrderBufs& b, uint32_t* rows,
                        uint32_t nr, const double* norms) {
  const double* nv = o.tani ? norms : nullptr;
  if (nr == 0 || nr > RESCAN_PRUNE_ROWS) return hipErrorInvalidValue;
  uint32_t p2 = 1;
  while (p2 < nr) p2 <<= 1;
  sort_rows_kernel<<<1, 1024, 0, s>>>(rows, nr, p2, b.perm, b.iperm);
  const uint32_t nblk = (nr + o.qb - 1) / o.qb;
  return by_dpm(o.dp, [&](auto dpm) -> hipError_t {
    constexpr int DPM = decltype(dpm)::value;
    cone_kernel<DPM><<<nblk, 64, 0, s>>>(xs32, o.dp, rows, nr, o.qb, b.rcen, b.rang, nv, b.rnr);
    bound_sort_kernel<<<nblk, 1024, 0, s>>>(b.rcen, b.rang, b.tcen, b.tang, o.dp, o.npt, b.rseq, b.rsb,
                                            nv ? b.rnr : nullptr, nv ? b.tnr : nullptr);
    return hipGetLastError();
  });
}

}  // namespace costal
