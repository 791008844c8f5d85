// mlp.hip -- the learned admission threshold: a 3-layer MLP d -> 64 -> 64 -> 1 on L2-normalized
// rows, trained on GPU by its own forward/backward/Adam kernels, run in BF16 (or FP32) to propose a
// per-row admission threshold for the fused kNN scan.
//
// Parameter layout (one flat f32 array, P = 64 d + 64 + 64*64 + 64 + 64 + 1 entries):
//   W1[64][d] (row-major), b1[64], W2[64][64] (row-major, out x in), b2[64], w3[64], b3
// Model: xn = f32(x / |x|) with |x| the f64 sequential norm (0 for zero rows),
//        h1 = relu(W1 xn + b1), h2 = relu(W2 h1 + b2), y = w3 . h2 + b3.
// Loss: mean over the batch of (y - target)^2.
//
// Everything here is deterministic: each parameter's gradient is one thread's fixed-order sum over
// the batch rows (no atomics), so a training run repeats bit for bit.

#include <hip/hip_runtime.h>

#include "device_ids.h"

#include <stdint.h>

#include <algorithm>

#include "../../include/costal_hip.h"

namespace mlp {

constexpr int H = COSTAL_MLP_HIDDEN;
constexpr int DMAX = COSTAL_MLP_MAX_FEATURES;
static_assert(DMAX == H, "the training kernel writes feature j from thread j");

__host__ __device__ constexpr uint64_t param_count(uint64_t d) { return H * d + H + H * H + H + H + 1; }

// round-to-nearest-even to bfloat16, kept in an f32
__device__ __forceinline__ float bf16r(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}

// Row features: the row scaled by its f64 sequential norm, rounded to f32 (the same working copy
// the kNN scan uses); zero rows give zeros.
__device__ __forceinline__ void load_features(const float* __restrict__ row, uint32_t d, float (&xn)[DMAX]) {
  double s = 0.0;
  for (uint32_t t = 0; t < d; ++t) s += (double)row[t] * (double)row[t];
  const double nrm = sqrt(s);
#pragma unroll
  for (int t = 0; t < DMAX; ++t) xn[t] = ((uint32_t)t < d && nrm > 0.0) ? (float)((double)row[t] / nrm) : 0.0f;
}

struct Smem {
  float W1[H * DMAX];
  float b1[H];
  float W2[H * H];
  float b2[H];
  float w3[H];
  float b3;
};

template <bool BF16>
__device__ void load_params(Smem& S, const float* __restrict__ P, uint32_t d) {
  const float* W1 = P;
  const float* b1 = W1 + H * d;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* w3 = b2 + H;
  const float* b3 = w3 + H;
  for (int e = tid_x(); e < H * DMAX; e += bdim_x()) {
    const int j = e / DMAX, t = e % DMAX;
    const float v = (uint32_t)t < d ? W1[j * d + t] : 0.0f;
    S.W1[e] = BF16 ? bf16r(v) : v;
  }
  for (int e = tid_x(); e < H * H; e += bdim_x()) S.W2[e] = BF16 ? bf16r(W2[e]) : W2[e];
  for (int e = tid_x(); e < H; e += bdim_x()) {
    S.b1[e] = b1[e];
    S.b2[e] = b2[e];
    S.w3[e] = BF16 ? bf16r(w3[e]) : w3[e];
  }
  if (tid_x() == 0) S.b3 = b3[0];
}

// Forward for one row in registers.  BF16: inputs, weights and hidden activations rounded to bf16,
// products accumulated in f32 (fmaf).
template <bool BF16>
__device__ __forceinline__ float forward_row(const Smem& S, float (&xn)[DMAX], float (&h1)[H], float (&h2)[H]) {
  if (BF16) {
#pragma unroll
    for (int t = 0; t < DMAX; ++t) xn[t] = bf16r(xn[t]);
  }
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float a = S.b1[j];
#pragma unroll
    for (int t = 0; t < DMAX; ++t) a = fmaf(S.W1[j * DMAX + t], xn[t], a);
    a = fmaxf(a, 0.0f);
    h1[j] = BF16 ? bf16r(a) : a;
  }
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float a = S.b2[j];
#pragma unroll
    for (int g = 0; g < H; ++g) a = fmaf(S.W2[j * H + g], h1[g], a);
    a = fmaxf(a, 0.0f);
    h2[j] = BF16 ? bf16r(a) : a;
  }
  float y = S.b3;
#pragma unroll
  for (int g = 0; g < H; ++g) y = fmaf(S.w3[g], h2[g], y);
  return y;
}

// Inference: out[r] = transform(y_r) where the model predicts u = log(1 - s + 1e-6) of the k-th
// neighbor similarity s; the threshold is 1 - exp(u + offset) - margin (offset and margin from the
// calibration), or y - offset when raw != 0.
template <bool BF16>
__global__ __launch_bounds__(256) void predict_kernel(const float* __restrict__ x, uint32_t n, uint32_t d,
                                                      const float* __restrict__ P, float offset, float margin,
                                                      int raw, float* __restrict__ out) {
  __shared__ Smem S;
  load_params<BF16>(S, P, d);
  __syncthreads();
  const uint32_t r = bid_x() * bdim_x() + tid_x();
  if (r >= n) return;
  float xn[DMAX], h1[H], h2[H];
  load_features(x + (size_t)r * d, d, xn);
  const float y = forward_row<BF16>(S, xn, h1, h2);
  out[r] = raw ? y - offset : 1.0f - expf(y + offset) - margin;
}

// BF16 inference on the matrix cores (v_mfma_f32_32x32x16_bf16): one wave per 32 rows at a time.
// The layers are computed transposed, hidden units x rows: H1^T = W1 Xn^T, H2^T = W2 H1^T.  In a
// 32x32 accumulator, lane l holds row l % 32 and units 8 (v / 4) + 4 (l / 32) + v % 4 of the tile
// (v = 0..15), so values 8c'..8c'+7 of the two lane halves together cover 16 units: layer 1's
// accumulators, after the bias, ReLU and BF16 rounding, are layer 2's B operand as they stand,
// with each K chunk in that permuted unit order, and W2's A operand is loaded in the same order.
// Inputs, weights and activations are rounded to BF16 as in forward_row<true>; their products are
// exact in f32 and accumulate in f32 in the matrix core's order (forward_row's fmaf chain agrees to
// f32 rounding).  A thread per row on the vector ALUs (predict_kernel) took 198 us for 396,460
// rows of 15 features, beside the kNN call's first kernels, which it slowed.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// the hidden unit of accumulator value v of lane half h, in 32-unit tile T
__device__ __forceinline__ int acc_unit(int T, int v, int h) { return 32 * T + 8 * (v >> 2) + 4 * h + (v & 3); }

// eight BF16-rounded f32 values as BF16 (element m in the low half of dword m / 2 when m is even)
__device__ __forceinline__ bf16x8 pack_bf16(const float (&v)[8]) {
  uint32_t u[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) u[q] = (__float_as_uint(v[2 * q]) >> 16) | (__float_as_uint(v[2 * q + 1]) & 0xffff0000u);
  return __builtin_bit_cast(bf16x8, make_uint4(u[0], u[1], u[2], u[3]));
}

// KC: 16-feature chunks of the first layer (d <= 16 KC)
template <int KC>
__global__ __launch_bounds__(256) void predict_mfma_kernel(const float* __restrict__ x, uint32_t n, uint32_t d,
                                                           const float* __restrict__ P, float offset, float margin,
                                                           int raw, float* __restrict__ out) {
  __shared__ float s_b1[H], s_b2[H], s_w3[H];
  const float* W1 = P;
  const float* b1 = W1 + H * d;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* w3 = b2 + H;
  const float b3 = w3[H];
  for (int e = tid_x(); e < H; e += bdim_x()) {
    s_b1[e] = b1[e];
    s_b2[e] = b2[e];
    s_w3[e] = bf16r(w3[e]);
  }
  __syncthreads();
  const int lane = tid_x() & 63, i = lane & 31, h = lane >> 5;
  // the weights in this lane's A-operand layout: row i of each 32-row tile, K elements 8h..8h+7
  bf16x8 w1op[2][KC], w2op[2][4];
#pragma unroll
  for (int T = 0; T < 2; ++T) {
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      float v[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const uint32_t f = 16 * c + 8 * h + m;
        v[m] = f < d ? bf16r(W1[(size_t)(32 * T + i) * d + f]) : 0.0f;
      }
      w1op[T][c] = pack_bf16(v);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float v[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = bf16r(W2[(32 * T + i) * H + acc_unit(c >> 1, 8 * (c & 1) + m, h)]);
      w2op[T][c] = pack_bf16(v);
    }
  }
  const uint32_t nwaves = gdim_x() * (bdim_x() >> 6);
  for (uint32_t r0 = (bid_x() * (bdim_x() >> 6) + (tid_x() >> 6)) * 32; r0 < n; r0 += nwaves * 32) {
    // this lane's row (both halves hold it): f64 sequential norm, features f32(x / norm) in BF16
    const uint32_t r = r0 + i < n ? r0 + i : n - 1;
    const float* row = x + (size_t)r * d;
    double s = 0.0;
    for (uint32_t t = 0; t < d; ++t) s += (double)row[t] * (double)row[t];
    const double nrm = sqrt(s);
    bf16x8 xop[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      float v[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const uint32_t f = 16 * c + 8 * h + m;
        v[m] = (f < d && nrm > 0.0) ? bf16r((float)((double)row[f] / nrm)) : 0.0f;
      }
      xop[c] = pack_bf16(v);
    }
    // layer 1, accumulators starting at the bias
    f32x16 a1[2];
#pragma unroll
    for (int T = 0; T < 2; ++T) {
#pragma unroll
      for (int v = 0; v < 16; ++v) a1[T][v] = s_b1[acc_unit(T, v, h)];
#pragma unroll
      for (int c = 0; c < KC; ++c) a1[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1op[T][c], xop[c], a1[T], 0, 0, 0);
    }
    bf16x8 hop[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float v[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = bf16r(fmaxf(a1[c >> 1][8 * (c & 1) + m], 0.0f));
      hop[c] = pack_bf16(v);
    }
    // layer 2, then the output unit over this lane's 32 hidden units and its partner lane's
    float y = 0.0f;
#pragma unroll
    for (int T = 0; T < 2; ++T) {
      f32x16 a2;
#pragma unroll
      for (int v = 0; v < 16; ++v) a2[v] = s_b2[acc_unit(T, v, h)];
#pragma unroll
      for (int c = 0; c < 4; ++c) a2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2op[T][c], hop[c], a2, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 16; ++v) y = fmaf(s_w3[acc_unit(T, v, h)], bf16r(fmaxf(a2[v], 0.0f)), y);
    }
    y = b3 + (y + __shfl_xor(y, 32));
    if (h == 0 && r0 + i < n) out[r0 + i] = raw ? y - offset : 1.0f - expf(y + offset) - margin;
  }
}

// Training forward + backward of the activations for one batch; per row it stores
// act = [xn (DMAX), h1 (H), h2 (H), dy, dh2 (H), dh1 (H), sq_err].
constexpr int ACT = DMAX + 5 * H + 2;

// One workgroup = one wave and FB_ROWS rows; thread j owns hidden unit j of both layers, so every
// dot product is one thread's fixed-order fmaf chain (the same order as forward_row) over weights
// held in LDS in the layout that makes the wave's reads consecutive: W1 transposed ([t][j]) and W2
// both ways ([j][g] for the backward pass, [g][j] for the forward pass).
constexpr int FB_ROWS = 8;

__global__ __launch_bounds__(64) void fwd_bwd_kernel(const float* __restrict__ x, const float* __restrict__ tgt,
                                                     uint32_t n, uint32_t d, const float* __restrict__ P,
                                                     float* __restrict__ act) {
  __shared__ float W1T[DMAX * H];
  __shared__ float W2R[H * H];
  __shared__ float W2T[H * H];
  __shared__ float xs[FB_ROWS][DMAX], h1s[FB_ROWS][H], h2s[FB_ROWS][H], dh2s[FB_ROWS][H], dys[FB_ROWS];
  const float* W1 = P;
  const float* b1 = W1 + H * d;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* w3 = b2 + H;
  const float* b3 = w3 + H;
  const uint32_t j = tid_x();
  for (uint32_t e = j; e < H * d; e += H) {
    const uint32_t jj = e / d, t = e % d;
    W1T[t * H + jj] = W1[e];
  }
  for (uint32_t e = j; e < H * H; e += H) {
    W2R[e] = W2[e];
    W2T[(e % H) * H + e / H] = W2[e];
  }
  const float b1j = b1[j], b2j = b2[j], w3j = w3[j];
  const uint32_t r0 = bid_x() * FB_ROWS;
  const uint32_t nr = n - r0 < (uint32_t)FB_ROWS ? n - r0 : (uint32_t)FB_ROWS;
  // features: thread rr computes row rr's f64 norm and writes its scaled row
  if (j < nr) {
    float xn[DMAX];
    load_features(x + (size_t)(r0 + j) * d, d, xn);
#pragma unroll
    for (int t = 0; t < DMAX; ++t) xs[j][t] = xn[t];
  }
  __syncthreads();
  for (uint32_t rr = 0; rr < nr; ++rr) {
    float a = b1j;
    for (uint32_t t = 0; t < d; ++t) a = fmaf(W1T[t * H + j], xs[rr][t], a);
    h1s[rr][j] = fmaxf(a, 0.0f);
  }
  __syncthreads();
  for (uint32_t rr = 0; rr < nr; ++rr) {
    float a = b2j;
#pragma unroll 8
    for (int g = 0; g < H; ++g) a = fmaf(W2T[g * H + j], h1s[rr][g], a);
    h2s[rr][j] = fmaxf(a, 0.0f);
  }
  __syncthreads();
  // output, error and dy: thread rr, a fixed-order sum over the hidden units
  if (j < nr) {
    float y = b3[0];
    for (int g = 0; g < H; ++g) y = fmaf(w3[g], h2s[j][g], y);
    const float err = y - tgt[r0 + j];
    dys[j] = 2.0f * err / (float)n;
    float* a = act + (size_t)(r0 + j) * ACT;
    a[DMAX + 2 * H] = dys[j];
    a[DMAX + 4 * H + 1] = err * err;
  }
  __syncthreads();
  for (uint32_t rr = 0; rr < nr; ++rr) dh2s[rr][j] = h2s[rr][j] > 0.0f ? w3j * dys[rr] : 0.0f;
  __syncthreads();
  for (uint32_t rr = 0; rr < nr; ++rr) {
    float* a = act + (size_t)(r0 + rr) * ACT;
    float s = 0.0f;
#pragma unroll 8
    for (int jj = 0; jj < H; ++jj) s = fmaf(W2R[jj * H + j], dh2s[rr][jj], s);
    a[j] = xs[rr][j];  // DMAX == H: thread j writes feature j
    a[DMAX + j] = h1s[rr][j];
    a[DMAX + H + j] = h2s[rr][j];
    a[DMAX + 2 * H + 1 + j] = dh2s[rr][j];
    a[DMAX + 3 * H + 1 + j] = h1s[rr][j] > 0.0f ? s : 0.0f;
  }
}

// Parameter gradients: thread p sums its parameter's contribution over the rows in order; thread P
// writes the batch loss (mean squared error).
__global__ __launch_bounds__(256) void wgrad_kernel(const float* __restrict__ act, uint32_t n, uint32_t d,
                                                    float* __restrict__ grad, float* __restrict__ loss) {
  const uint32_t P = (uint32_t)param_count(d);
  const uint32_t p = bid_x() * bdim_x() + tid_x();
  if (p > P) return;
  // offsets inside a row of act
  const int XN = 0, H1 = DMAX, H2 = DMAX + H, DY = DMAX + 2 * H, DH2 = DY + 1, DH1 = DH2 + H, SQ = DH1 + H;
  int ia, ib;  // grad = sum_r act[r][ia] * (ib < 0 ? 1 : act[r][ib])
  if (p == P) {
    ia = SQ;
    ib = -1;
  } else if (p < H * d) {
    ia = DH1 + (int)(p / d);
    ib = XN + (int)(p % d);
  } else if (p < H * d + H) {
    ia = DH1 + (int)(p - H * d);
    ib = -1;
  } else if (p < H * d + H + H * H) {
    const uint32_t q = p - (H * d + H);
    ia = DH2 + (int)(q / H);
    ib = H1 + (int)(q % H);
  } else if (p < H * d + 2 * H + H * H) {
    ia = DH2 + (int)(p - (H * d + H + H * H));
    ib = -1;
  } else if (p < H * d + 3 * H + H * H) {
    ia = DY;
    ib = H2 + (int)(p - (H * d + 2 * H + H * H));
  } else {
    ia = DY;
    ib = -1;
  }
  float s = 0.0f;
  for (uint32_t r = 0; r < n; ++r) {
    const float* a = act + (size_t)r * ACT;
    s = fmaf(a[ia], ib < 0 ? 1.0f : a[ib], s);
  }
  if (p == P) {
    if (loss) loss[0] = s / (float)n;
  } else {
    grad[p] = s;
  }
}

__global__ void adam_kernel(uint32_t count, float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const uint32_t i = bid_x() * bdim_x() + tid_x();
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

thread_local char g_err[256];

int fail(int code, const char* msg, const char* detail = "") {
  snprintf(g_err, sizeof g_err, "%s%s", msg, detail);
  return code;
}

}  // namespace mlp

extern "C" {

const char* costal_mlp_last_error(void) { return mlp::g_err; }

uint64_t costal_mlp_param_count(uint64_t d) { return mlp::param_count(d); }

uint64_t costal_mlp_act_floats(uint64_t n) { return n * (uint64_t)mlp::ACT; }

int costal_mlp_predict_device(const float* x, uint64_t n, uint64_t d, const float* params, uint32_t bf16,
                              float offset, float margin, uint32_t raw, float* out, void* stream) {
  if (d == 0 || d > COSTAL_MLP_MAX_FEATURES) return mlp::fail(COSTAL_E_UNSUPPORTED, "the threshold model takes 1..64 features");
  if (n == 0) return 0;
  if (n > 0xffffffffull) return mlp::fail(COSTAL_E_TOO_MANY_ROWS, "too many rows");
  hipStream_t s = (hipStream_t)stream;
  if (bf16) {
    // four waves of 32-row batches per workgroup; the waves loop when there are more batches
    // than 1,024 workgroups hold (each wave loads the weights once)
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 127) / 128, 1024);
    const uint32_t kc = (uint32_t)((d + 15) / 16);
#define PREDICT(KC) \
  mlp::predict_mfma_kernel<KC><<<grid, 256, 0, s>>>(x, (uint32_t)n, (uint32_t)d, params, offset, margin, (int)raw, out)
    if (kc == 1) PREDICT(1);
    else if (kc == 2) PREDICT(2);
    else if (kc == 3) PREDICT(3);
    else PREDICT(4);
#undef PREDICT
  } else {
    const unsigned grid = (unsigned)((n + 255) / 256);
    mlp::predict_kernel<false><<<grid, 256, 0, s>>>(x, (uint32_t)n, (uint32_t)d, params, offset, margin, (int)raw, out);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mlp::fail(COSTAL_E_DEVICE, "predict launch failed: ", hipGetErrorString(e));
  return 0;
}

int costal_mlp_grad_device(const float* x, const float* target, uint64_t n, uint64_t d, const float* params,
                           float* act, float* grad, float* loss, void* stream) {
  if (d == 0 || d > COSTAL_MLP_MAX_FEATURES) return mlp::fail(COSTAL_E_UNSUPPORTED, "the threshold model takes 1..64 features");
  if (n == 0 || n > 0xffffffffull) return mlp::fail(COSTAL_E_INVALID, "bad batch size");
  hipStream_t s = (hipStream_t)stream;
  mlp::fwd_bwd_kernel<<<(unsigned)((n + mlp::FB_ROWS - 1) / mlp::FB_ROWS), 64, 0, s>>>(x, target, (uint32_t)n, (uint32_t)d, params, act);
  const uint64_t P = mlp::param_count(d);
  mlp::wgrad_kernel<<<(unsigned)((P + 1 + 255) / 256), 256, 0, s>>>(act, (uint32_t)n, (uint32_t)d, grad, loss);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mlp::fail(COSTAL_E_DEVICE, "gradient launch failed: ", hipGetErrorString(e));
  return 0;
}

int costal_mlp_adam_device(uint64_t count, float* params, const float* grad, float* m, float* v, float lr,
                           float beta1, float beta2, float eps, uint64_t step, void* stream) {
  if (step == 0) return mlp::fail(COSTAL_E_INVALID, "Adam steps count from 1");
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  mlp::adam_kernel<<<(unsigned)((count + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      (uint32_t)count, params, grad, m, v, lr, beta1, beta2, eps, c1, c2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mlp::fail(COSTAL_E_DEVICE, "adam launch failed: ", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
