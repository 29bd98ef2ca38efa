// k_closure.hip -- the learned eddy-viscosity closure: a 10 -> 64 -> 64 -> 1
// multilayer perceptron (ReLU hidden layers, softplus output) with BF16
// weights and FP32 accumulation on the MFMA units of gfx950, its training
// (forward, backward, Adam) and its inference.  The reference has no closure.
//
// Tiles.  A wave owns 32 samples at a time.  The hidden layers are
// v_mfma_f32_32x32x16_bf16 products: A is 32 x 16 (lane l holds row l%32,
// k = 8*(l/32) .. +7), B is 16 x 32 (lane l holds column l%32, the same k),
// D is 32 x 32 FP32 (lane l holds column l%32, rows 8*(j/4) + 4*(l/32) + j%4
// for j = 0..15).  In training and batch inference activations pass between
// layers through LDS as BF16 (the A operand of the next product), the output
// layer is one more product with W3 in every B column, and the weight
// gradients take both operands from the accumulator registers.  The closure
// stage computes the layers transposed, entirely in registers, with the
// output layer as BF16 dot products.  Rounding to BF16 is round-to-nearest-
// even at fixed points -- the inputs, the weights, the activations h1 and h2
// and the back-propagated deltas -- so a CPU restatement can follow it
// exactly; only the FP32 summation order differs.
//
// Training is deterministic: each wave accumulates its gradient over its
// tiles in registers, a block sums its waves' partials in LDS in wave order
// and writes one partial, and two more levels sum the partials in a fixed
// order, the last inside the Adam update.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <random>
#include <vector>

#include "pll_internal.hpp"
#include "stencil_ops.hpp"

namespace pll {
namespace {

constexpr int kIn = 10;   // features
constexpr int kInP = 16;  // padded to one MFMA K block
constexpr int kH = 64;    // hidden width
// parameter vector (FP32 master): W1[16][64], b1[64], W2[64][64], b2[64], W3[64], b3
constexpr int oW1 = 0, oB1 = oW1 + kInP * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH, kNParam = oB3 + 1;
constexpr int kWaves = 4;          // waves per block
constexpr long kMaxWaves = 1024;   // training grid: one wave per SIMD (one partial per kWaves waves)
constexpr int kRedChunk = 32;      // block partials summed per chunk in the first reduction level

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// FP32 -> BF16, round to nearest, ties to even (gfx950 v_cvt_pk_bf16_f32).
__device__ inline uint16_t to_bf16(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
__host__ __device__ inline float from_bf16(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  memcpy(&x, &u, 4);
  return x;
}

// softplus(o) = log(1 + e^o) = max(o, 0) + log1p(t), t = e^-|o| in (0, 1],
// with log1p(t) in Goldberg's form log(u) t / (u - 1), u = 1 + t rounded,
// which keeps the relative accuracy of the logarithm for small t (D. Goldberg,
// "What every computer scientist should know about floating-point
// arithmetic", 1991, Theorem 4); the exponential, logarithm and quotient on the
// hardware exp2/log2/reciprocal (relative error ~1e-6 in all; u lies in
// [1, 2], so no denormal scaling is needed).
__device__ inline float softplus(float o) {
  constexpr float kLog2e = 1.44269504088896341f, kLn2 = 0.693147180559945309f;
  const float t = __builtin_amdgcn_exp2f(-fabsf(o) * kLog2e);
  const float u = 1.f + t;
  const float l = u == 1.f ? t : (__builtin_amdgcn_logf(u) * kLn2) * (t * __builtin_amdgcn_rcpf(u - 1.f));
  return fmaxf(o, 0.f) + l;
}
__device__ inline float sigmoid(float o) { return 1.f / (1.f + expf(-o)); }

struct Mlp {
  const float* p;      // FP32 parameters (biases and output weights are read from here)
  const uint16_t* w1;  // BF16 W1 [16][64]
  const uint16_t* w2;  // BF16 W2 [64][64]
  const uint16_t* w3;  // BF16 W3 [64]
  // the same BF16 weights as each lane's B operand of the forward products,
  // 8 values (16 bytes) per lane: b1[t][lane][e] = W1[8h + e][32t + c] and
  // b2[t][kb][lane][e] = W2[16kb + 8h + e][32t + c], h = lane/32, c = lane%32
  const uint16_t* b1;
  const uint16_t* b2;
};

// Layout of the BF16 weight buffer: W1 | W2 | W3 | b1 | b2 (element offsets).
constexpr int kWb1 = 0, kWb2 = kWb1 + kInP * kH, kWb3 = kWb2 + kH * kH, kWbB1 = kWb3 + kH,
              kWbB2 = kWbB1 + 2 * 64 * 8, kWbLen = kWbB2 + 2 * 4 * 64 * 8;
static_assert(kWbB1 % 8 == 0 && kWbB2 % 8 == 0, "operand copies start on 16-byte boundaries");

// BF16 of parameter i (already computed as v) into the plain and operand copies.
__device__ inline void store_bf16_weight(uint16_t* __restrict__ wb, int i, float v) {
  const uint16_t b = to_bf16(v);
  if (i < oB1) {
    const int r = (i - oW1) / kH, c = (i - oW1) % kH;
    wb[kWb1 + i - oW1] = b;
    const int t = c / 32, lane = 32 * ((r % 16) / 8) + c % 32;
    wb[kWbB1 + (t * 64 + lane) * 8 + r % 8] = b;
  } else if (i >= oW2 && i < oB2) {
    const int r = (i - oW2) / kH, c = (i - oW2) % kH;
    wb[kWb2 + i - oW2] = b;
    const int t = c / 32, kb = r / 16, lane = 32 * ((r % 16) / 8) + c % 32;
    wb[kWbB2 + ((t * 4 + kb) * 64 + lane) * 8 + r % 8] = b;
  } else if (i >= oW3 && i < oB3) {
    wb[kWb3 + i - oW3] = b;
  }
}

// Orders a wave's LDS writes before its lanes' later LDS reads of them.  A
// wave's LDS instructions execute in order, so this only has to keep the
// compiler from moving memory operations across it: the training and predict
// kernels share no LDS between waves and need no block barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wave LDS: X (32 x 16), h1, h2 and delta2 (32 x 64), all BF16.
struct WaveLds {
  uint16_t x[32][kInP];
  uint16_t h1[32][kH];
  uint16_t h2[32][kH];
  uint16_t d2[32][kH];
  float o[32];    // network output per sample
  float g[32];    // dL/do per sample
  float red[64];  // lane exchange
};

__device__ inline bf16x8 load8(const uint16_t* p) {
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (short)p[e];
  return v;
}

// BF16 (nearest even) of accumulator elements j0 .. j0+7 as one MFMA operand.
__device__ inline bf16x8 pack8_bf16(const f32x16& v, int j0) {
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; ++e) r[e] = (short)to_bf16(v[j0 + e]);
  return r;
}

// D-layout row of accumulator element j for this lane.
__device__ inline int drow(int j, int lane) { return 8 * (j / 4) + 4 * (lane / 32) + (j % 4); }

// X of one 32-sample tile from a feature array (rows beyond n are zero).
template <class L_>
__device__ inline void stage_x(const float* __restrict__ x, long m0, long n, L_& L, int lane) {
  for (int e = lane; e < 32 * kInP; e += 64) {
    const int m = e / kInP, k = e % kInP;
    const long s = m0 + m;
    L.x[m][k] = to_bf16((k < kIn && s < n) ? x[s * kIn + k] : 0.f);
  }
}

// Forward pass of one 32-sample tile whose BF16 features are in L.x (written
// before the call by any lane of the wave): fills L.h1/h2, returns the ReLU
// activations in D layout and, on lanes 0..31, the output o of sample lane.
// Only the calling wave's LDS is used.
template <class L_>
__device__ inline void forward_tile(const Mlp& w, L_& L, int lane, f32x16 h1[2], f32x16 h2[2],
                                    float& out_o) {
  wave_lds_sync();
  const int col = lane % 32, kh = 8 * (lane / 32);
  // layer 1: h1 = relu(X W1 + b1)
  const bf16x8 a1 = load8(&L.x[col][kh]);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(w.b1 + (t * 64 + lane) * 8);
    f32x16 acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
    const float bias = w.p[oB1 + 32 * t + col];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float h = fmaxf(acc[j] + bias, 0.f);
      acc[j] = h;
      L.h1[drow(j, lane)][32 * t + col] = to_bf16(h);
    }
    h1[t] = acc;
  }
  wave_lds_sync();
  // layer 2: h2 = relu(h1 W2 + b2)
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    f32x16 acc = {};
#pragma unroll
    for (int kb = 0; kb < kH / 16; ++kb) {
      const bf16x8 a = load8(&L.h1[col][16 * kb + kh]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(w.b2 + ((t * 4 + kb) * 64 + lane) * 8);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const float bias = w.p[oB2 + 32 * t + col];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float h = fmaxf(acc[j] + bias, 0.f);
      acc[j] = h;
      L.h2[drow(j, lane)][32 * t + col] = to_bf16(h);
    }
    h2[t] = acc;
  }
  wave_lds_sync();
  // output: o = h2 . W3 + b3 as one more product, B holding W3 in all 32
  // columns (K = hidden units); column 0 goes back to lanes 0..31 through LDS
  f32x16 acc = {};
#pragma unroll
  for (int kb = 0; kb < kH / 16; ++kb) {
    const bf16x8 a = load8(&L.h2[col][16 * kb + kh]);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(w.w3 + 16 * kb + kh);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  if (col == 0)
#pragma unroll
    for (int j = 0; j < 16; ++j) L.o[drow(j, lane)] = acc[j];
  wave_lds_sync();
  out_o = lane < 32 ? L.o[lane] + w.p[oB3] : 0.f;
}

// Training: one partial gradient (and loss sum) per block.  Loss = mean over
// the n samples of (softplus(o) - y)^2.
template <bool TRAIN>
__global__ __launch_bounds__(64 * kWaves) void k_mlp(Mlp w, const float* __restrict__ x,
                                                      const float* __restrict__ y, long n,
                                                      float* __restrict__ out,
                                                      float* __restrict__ partial,
                                                      int tiles_per_wave) {
  __shared__ WaveLds lds[kWaves];
  __shared__ float bp[TRAIN ? kWaves : 1][kNParam + 1];  // the waves' partial gradients
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  WaveLds& L = lds[wave];
  const long gw = (long)blockIdx.x * kWaves + wave;  // global wave index
  const int col = lane % 32, kh = 8 * (lane / 32);
  f32x16 gW2[2][2] = {};  // dW2 tiles [k-tile][n-tile] (rows k, cols n)
  f32x16 gW1[2] = {};     // dW1 tiles [n-tile] (rows i < 16 meaningful)
  float gb1[2] = {0.f, 0.f}, gb2[2] = {0.f, 0.f}, gw3[2] = {0.f, 0.f}, gb3 = 0.f, loss = 0.f;
  const float inv_n = 1.f / (float)n;
  const long nwaves = (long)gridDim.x * kWaves;
  for (int tile = 0; tile < tiles_per_wave; ++tile) {
    const long m0 = (gw + tile * nwaves) * 32;
    f32x16 h1[2], h2[2];
    float o;
    stage_x(x, m0, n, L, lane);
    forward_tile(w, L, lane, h1, h2, o);
    if (!TRAIN) {
      if (lane < 32 && m0 + lane < n) out[m0 + lane] = softplus(o);
      wave_lds_sync();
      continue;
    }
    // dL/do
    if (lane < 32) {
      const long s = m0 + lane;
      float g = 0.f;
      if (s < n) {
        const float d = softplus(o) - y[s];
        loss += d * d;
        g = 2.f * d * inv_n * sigmoid(o);
      }
      L.g[lane] = g;
    }
    wave_lds_sync();
    // output layer gradients and delta2 = g w3 relu'(z2), in the D layout
    // (the lane's column nn, samples drow(j, lane))
    f32x16 d2r[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int nn = 32 * t + col;
      const float w3n = from_bf16(w.w3[nn]);
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int m = drow(j, lane);
        const float g = L.g[m];
        const float h = h2[t][j];
        gw3[t] += from_bf16(to_bf16(h)) * g;
        const float d2 = h > 0.f ? g * w3n : 0.f;
        gb2[t] += d2;
        d2r[t][j] = d2;
        L.d2[m][nn] = to_bf16(d2);
      }
    }
    if (lane == 0)
      for (int m = 0; m < 32; ++m) gb3 += L.g[m];
    wave_lds_sync();
    // dW2 += h1^T d2 : rows k (h1 column), cols n, K = samples.  Both operands
    // come from the D-layout registers: K-slot e of the lane's half in block ks
    // is sample drow(8 ks + e, lane), the same in A and B.
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        a[t] = pack8_bf16(h1[t], 8 * ks);
        b[t] = pack8_bf16(d2r[t], 8 * ks);
      }
#pragma unroll
      for (int tk = 0; tk < 2; ++tk)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
          gW2[tk][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[tk], b[tn], gW2[tk][tn], 0, 0, 0);
    }
    // delta1 = (d2 W2^T) relu'(z1) : rows m, cols k, K = n
    f32x16 d1r[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      f32x16 acc = {};
#pragma unroll
      for (int kb = 0; kb < kH / 16; ++kb) {
        const bf16x8 a = load8(&L.d2[col][16 * kb + kh]);
        const bf16x8 b = load8(&w.w2[(32 * t + col) * kH + 16 * kb + kh]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const float h = h1[t][j];
        const float d1 = h > 0.f ? acc[j] : 0.f;
        gb1[t] += d1;
        acc[j] = d1;
      }
      d1r[t] = acc;
    }
    // dW1 += X^T d1 : rows i (< 16), cols k, K = samples in the order of dW2
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        a[e] = col < kInP ? (short)L.x[drow(8 * ks + e, lane)][col] : (short)0;
#pragma unroll
      for (int tn = 0; tn < 2; ++tn)
        gW1[tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, pack8_bf16(d1r[tn], 8 * ks), gW1[tn],
                                                          0, 0, 0);
    }
    wave_lds_sync();
  }
  if (!TRAIN) return;
  // this wave's partial into LDS: lanes l and l+32 hold different rows of the
  // same columns; the per-column sums (biases, W3) are combined through LDS
  float* P = bp[wave];
#pragma unroll
  for (int tk = 0; tk < 2; ++tk)
#pragma unroll
    for (int tn = 0; tn < 2; ++tn)
#pragma unroll
      for (int j = 0; j < 16; ++j)
        P[oW2 + (32 * tk + drow(j, lane)) * kH + 32 * tn + col] = gW2[tk][tn][j];
#pragma unroll
  for (int tn = 0; tn < 2; ++tn)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int i = drow(j, lane);
      if (i < kInP) P[oW1 + i * kH + 32 * tn + col] = gW1[tn][j];
    }
  auto colsum = [&](float v, int t, int off) {
    L.red[lane] = v;
    wave_lds_sync();
    if (lane < 32) P[off + 32 * t + lane] = L.red[lane] + L.red[lane + 32];
    wave_lds_sync();
  };
  for (int t = 0; t < 2; ++t) {
    colsum(gb1[t], t, oB1);
    colsum(gb2[t], t, oB2);
    colsum(gw3[t], t, oW3);
  }
  L.red[lane] = lane < 32 ? loss : 0.f;
  wave_lds_sync();
  if (lane == 0) {
    float l = 0.f;
    for (int q = 0; q < 32; ++q) l += L.red[q];
    P[oB3] = gb3;
    P[kNParam] = l;
  }
  // the block's partial: its waves' partials summed in wave order
  __syncthreads();
  float* Q = partial + (long)blockIdx.x * (kNParam + 1);
  for (int i = threadIdx.x; i <= kNParam; i += 64 * kWaves) {
    float v = bp[0][i];
#pragma unroll
    for (int q = 1; q < kWaves; ++q) v += bp[q][i];
    Q[i] = v;
  }
}

// Sum of the per-block partials in a fixed order, in two levels: chunk c of
// kRedChunk partials into chunks[c], then the chunks in order into grad
// (grad[i] for i < kNParam, the loss sum last).
__global__ __launch_bounds__(256) void k_grad_reduce(const float* __restrict__ partial,
                                                     long nparts, float* __restrict__ chunks) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const long c = blockIdx.y;
  if (i > kNParam) return;
  const long q1 = std::min(nparts, (c + 1) * kRedChunk);
  float s = 0.f;
  for (long q = c * kRedChunk; q < q1; ++q) s += partial[q * (kNParam + 1) + i];
  chunks[c * (kNParam + 1) + i] = s;
}
__global__ __launch_bounds__(256) void k_grad_final(const float* __restrict__ chunks, long nchunk,
                                                    float* __restrict__ grad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i > kNParam) return;
  float s = 0.f;
  for (long c = 0; c < nchunk; ++c) s += chunks[c * (kNParam + 1) + i];
  grad[i] = s;
}

// Adam (Kingma & Ba, 2015) on the FP32 master parameters; refreshes the BF16
// weights.  The padded rows 10..15 of W1 stay zero.  The gradient is the
// chunk sums of k_grad_reduce added in k_grad_final's order (so the same
// bits), also stored in grad with the loss sum.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, float* __restrict__ m1,
                                              float* __restrict__ m2,
                                              const float* __restrict__ chunks, long nchunk,
                                              float* __restrict__ grad, float lr, float b1,
                                              float b2, float eps, float c1, float c2,
                                              uint16_t* __restrict__ wb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i > kNParam) return;
  float gs = 0.f;
  for (long c = 0; c < nchunk; ++c) gs += chunks[c * (kNParam + 1) + i];
  grad[i] = gs;
  if (i == kNParam) return;  // the loss sum
  const bool pad = i < oB1 && (i / kH) >= kIn;
  float v = p[i];
  if (!pad) {
    const float gi = gs;
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    v -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
  }
  store_bf16_weight(wb, i, pad ? 0.f : v);
}

__global__ void k_refresh_bf16(const float* __restrict__ p, uint16_t* __restrict__ wb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kNParam) return;
  store_bf16_weight(wb, i, (i < oB1 && (i / kH) >= kIn) ? 0.f : p[i]);
}

// ---------------------------------------------------------------------------
// Features, training samples and the closure stage.
//
// The nine velocity gradients at the centre of cell (i,j,k) of a staggered
// field (u at i+1/2, v at j+1/2, w at k+1/2), from its halo-padded arrays:
// g = {du/dx, du/dy, du/dz, dv/dx, dv/dy, dv/dz, dw/dx, dw/dy, dw/dz}.  Normal
// derivatives are the face differences across the cell; cross derivatives are
// central differences of the cell-centred averages of the neighbours
// (qdxi = dxi/4, qdyi = dyi/4, hz = 1/(2 (zc(k+1) - zc(k-1)))).  a = idx(i,j,k)
// with i, j interior and 1 <= k <= nz.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void cell_gradient(const T* __restrict__ u, const T* __restrict__ v,
                                     const T* __restrict__ w, const Geom& g, long a, T dxi,
                                     T dyi, T qdxi, T qdyi, T dfi, T hz, T gr[9]) {
  const long sx = g.sx, sz = g.sxy;
  gr[0] = (u[a] - u[a - 1]) * dxi;
  gr[1] = ((u[a + sx] + u[a - 1 + sx]) - (u[a - sx] + u[a - 1 - sx])) * qdyi;
  gr[2] = ((u[a + sz] + u[a - 1 + sz]) - (u[a - sz] + u[a - 1 - sz])) * hz;
  gr[3] = ((v[a + 1] + v[a + 1 - sx]) - (v[a - 1] + v[a - 1 - sx])) * qdxi;
  gr[4] = (v[a] - v[a - sx]) * dyi;
  gr[5] = ((v[a + sz] + v[a - sx + sz]) - (v[a - sz] + v[a - sx - sz])) * hz;
  gr[6] = ((w[a + 1] + w[a + 1 - sz]) - (w[a - 1] + w[a - 1 - sz])) * qdxi;
  gr[7] = ((w[a + sx] + w[a + sx - sz]) - (w[a - sx] + w[a - sx - sz])) * qdyi;
  gr[8] = (w[a] - w[a - sz]) * dfi;
}

// Network inputs: the gradients times gs, with the four components that
// change sign under the reflection z -> lz - z (du/dz, dv/dz, dw/dx, dw/dy)
// multiplied by zs = -1 in the upper half, so both walls look like the lower
// one; the distance to the nearer wall times ys.
template <typename T>
__device__ inline void features(const T gr[9], T gs, T ys, T yw, T zs, float f[kIn]) {
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    T q = gr[c] * gs;
    if (c == 2 || c == 5 || c == 6 || c == 7) q = q * zs;
    f[c] = (float)q;
  }
  f[9] = (float)(yw * ys);
}

// Per-plane coefficients of the features (length nz+2 each): hz, the wall
// distance of the cell centre and the reflection sign.
template <typename T>
struct FeatMesh {
  const T* hz;
  const T* yw;
  const T* zs;
};

// ---- box filter: the coarse staggered field whose every face value is the
// area average of the fine face values on it (fine faces nest in coarse ones:
// nx_f = fx nx_c, ny_f = fy ny_c, nz_f = fz nz_c with nesting z faces).  The
// coarse flux through each face is the sum of the fine fluxes, so a discretely
// divergence-free fine field gives a divergence-free coarse one.
template <typename Tf, typename Tc>
__global__ __launch_bounds__(256) void k_filter(Geom gf, const Tf* __restrict__ uf,
                                                const Tf* __restrict__ vf,
                                                const Tf* __restrict__ wf,
                                                const double* __restrict__ dzf_f, Geom gc,
                                                Tc* __restrict__ uc, Tc* __restrict__ vc,
                                                Tc* __restrict__ wc, int fx, int fy, int fz) {
  const long n = (long)blockIdx.x * 256 + threadIdx.x;
  const long ncell = (long)gc.nx * gc.ny * gc.nz;
  if (n >= ncell) return;
  const int I = (int)(n % gc.nx) + 1, J = (int)((n / gc.nx) % gc.ny) + 1,
            K = (int)(n / ((long)gc.nx * gc.ny)) + 1;
  const int ib = (I - 1) * fx, jb = (J - 1) * fy, kb = (K - 1) * fz;  // fine cells ib+1..ib+fx
  double su = 0.0, sv = 0.0, dzs = 0.0;
  for (int kk = 1; kk <= fz; ++kk) {
    const double dz = dzf_f[kb + kk];
    double tu = 0.0, tv = 0.0;
    for (int jj = 1; jj <= fy; ++jj) tu += (double)uf[gf.idx(I * fx, jb + jj, kb + kk)];
    for (int ii = 1; ii <= fx; ++ii) tv += (double)vf[gf.idx(ib + ii, J * fy, kb + kk)];
    su += tu * dz;
    sv += tv * dz;
    dzs += dz;
  }
  double sw = 0.0;
  for (int jj = 1; jj <= fy; ++jj)
    for (int ii = 1; ii <= fx; ++ii) sw += (double)wf[gf.idx(ib + ii, jb + jj, K * fz)];
  const long a = gc.idx(I, J, K);
  uc[a] = (Tc)(su / (fy * dzs));
  vc[a] = (Tc)(sv / (fx * dzs));
  wc[a] = (Tc)(sw / (fx * fy));
}

// ---- training samples: for every coarse cell, the features of the filtered
// field (the coarse context, which holds the filter of the fine one with its
// halos) and the label: the eddy viscosity nu* that best fits the subgrid
// stress of the fine field by least squares,
//   tau_ij = <uc_i uc_j> - <uc_i><uc_j>   (<.> the dz-weighted box average of
//                                          fine cell-centre velocities),
//   nu* = max(0, -tau^d_ij S_ij) / (2 S_ij S_ij + eps),
// with S the strain rate of the filtered field at the cell centre; the label
// is min(nu* / nus, ymax).
template <typename Tf, typename Tc>
__global__ __launch_bounds__(256) void k_samples(
    Geom gf, const Tf* __restrict__ uf, const Tf* __restrict__ vf, const Tf* __restrict__ wf,
    const double* __restrict__ dzf_f, Geom gc, const Tc* __restrict__ uc,
    const Tc* __restrict__ vc, const Tc* __restrict__ wc, const Tc* __restrict__ dzf_inv_c,
    FeatMesh<Tc> fm, Tc dxi, Tc dyi, Tc qdxi, Tc qdyi, Tc gs, Tc ys, double nus, double eps,
    double ymax, int fx, int fy, int fz, float* __restrict__ x, float* __restrict__ y) {
  const long n = (long)blockIdx.x * 256 + threadIdx.x;
  const long ncell = (long)gc.nx * gc.ny * gc.nz;
  if (n >= ncell) return;
  const int I = (int)(n % gc.nx) + 1, J = (int)((n / gc.nx) % gc.ny) + 1,
            K = (int)(n / ((long)gc.nx * gc.ny)) + 1;
  Tc gr[9];
  cell_gradient(uc, vc, wc, gc, gc.idx(I, J, K), dxi, dyi, qdxi, qdyi, dzf_inv_c[K], fm.hz[K], gr);
  float f[kIn];
  features(gr, gs, ys, fm.yw[K], fm.zs[K], f);
#pragma unroll
  for (int c = 0; c < kIn; ++c) x[n * kIn + c] = f[c];
  // box moments of the fine cell-centre velocities
  const int ib = (I - 1) * fx, jb = (J - 1) * fy, kb = (K - 1) * fz;
  double m[3] = {0, 0, 0}, q[6] = {0, 0, 0, 0, 0, 0}, wsum = 0.0;
  for (int kk = 1; kk <= fz; ++kk) {
    const double dz = dzf_f[kb + kk];
    double pm[3] = {0, 0, 0}, pq[6] = {0, 0, 0, 0, 0, 0};
    for (int jj = 1; jj <= fy; ++jj)
      for (int ii = 1; ii <= fx; ++ii) {
        const long a = gf.idx(ib + ii, jb + jj, kb + kk);
        const double cu = 0.5 * ((double)uf[a - 1] + (double)uf[a]);
        const double cv = 0.5 * ((double)vf[a - gf.sx] + (double)vf[a]);
        const double cw = 0.5 * ((double)wf[a - gf.sxy] + (double)wf[a]);
        pm[0] += cu;
        pm[1] += cv;
        pm[2] += cw;
        pq[0] += cu * cu;
        pq[1] += cv * cv;
        pq[2] += cw * cw;
        pq[3] += cu * cv;
        pq[4] += cu * cw;
        pq[5] += cv * cw;
      }
    for (int c = 0; c < 3; ++c) m[c] += pm[c] * dz;
    for (int c = 0; c < 6; ++c) q[c] += pq[c] * dz;
    wsum += dz * (fx * fy);
  }
  for (int c = 0; c < 3; ++c) m[c] /= wsum;
  for (int c = 0; c < 6; ++c) q[c] /= wsum;
  const double txx = q[0] - m[0] * m[0], tyy = q[1] - m[1] * m[1], tzz = q[2] - m[2] * m[2];
  const double txy = q[3] - m[0] * m[1], txz = q[4] - m[0] * m[2], tyz = q[5] - m[1] * m[2];
  const double sxx = gr[0], syy = gr[4], szz = gr[8];
  const double sxy = 0.5 * ((double)gr[1] + (double)gr[3]);
  const double sxz = 0.5 * ((double)gr[2] + (double)gr[6]);
  const double syz = 0.5 * ((double)gr[5] + (double)gr[7]);
  const double ss = ((sxx * sxx + syy * syy) + szz * szz) + 2.0 * ((sxy * sxy + sxz * sxz) + syz * syz);
  const double ts = ((txx * sxx + tyy * syy) + tzz * szz) + 2.0 * ((txy * sxy + txz * sxz) + tyz * syz) -
                    ((txx + tyy) + tzz) / 3.0 * ((sxx + syy) + szz);
  const double nu = fmax(0.0, -ts) / (2.0 * ss + eps);
  y[n] = (float)fmin(nu / nus, ymax);
}

// ---- the deferred projection, forcing and boundary update -----------------
// In the closure's step sequence the previous step's velocity has neither its
// projection nor its forcing nor its boundary update: its values are rebuilt
// from the stored ones as k_rk's folded loads do.  With `bcload`: x/y halo
// cells read their periodic image, the ghost planes 0 and nz+1 follow the
// boundary rules from planes 1 and nz (2 bc - u, -v;
// src/mod_updateBound.cc:92,123) and w is 0 on planes 0, nz and nz+1 (:137);
// with `corr` the value first becomes u - dt grad(p) (src/mod_calcVel.cc:272-274,
// k_correct's association), then the deferred forcing of each component in
// fmask (device scalars force[0..2]) is added.  Without `bcload` the stored
// value is read.  p's x halos must hold their periodic values (the step's
// inverse x transform writes them); its y halo is read at its image and its z
// ghost planes are not needed (w is 0 on the planes that would read them), so
// updateBoundP need not have run.
template <typename T>
struct VelRules {
  int bcload, corr;
  const T* p;              // previous pressure, valid halos (corr)
  const T* dzc_inv;        // (corr)
  T dtdxi, dtdyi, dt;      // (corr)
  const double* force;     // deferred forcing of u, v, w (fmask)
  int fmask;
  T bc_u;                  // u's wall value, src/mod_updateBound.cc:80
};

// component fc (0 u, 1 v, 2 w) of field q at (i, j, k), 0 <= i <= nx+1 etc.
template <typename T>
__device__ __forceinline__ T rule_load(const T* __restrict__ q, int fc, int i, int j, int k,
                                       const Geom& g, const VelRules<T>& r) {
  if (!r.bcload) return q[g.idx(i, j, k)];
  const int ii = i == 0 ? g.nx : (i == g.nx + 1 ? 1 : i);
  const int jj = j == 0 ? g.ny : (j == g.ny + 1 ? 1 : j);
  const bool ghost = k == 0 || k == g.nz + 1;
  const int ks = ghost ? (k == 0 ? 1 : g.nz) : k;
  if (fc == 2 && (k == 0 || k >= g.nz)) return T(0);
  const long a = g.idx(ii, jj, ks);
  T v = q[a];
  if (r.corr) {
    const T p0 = r.p[a];
    if (fc == 0)
      v = v - (r.p[a + 1] - p0) * r.dtdxi;
    else if (fc == 1)
      v = v - (r.p[jj == g.ny ? a - (long)(g.ny - 1) * g.sx : a + g.sx] - p0) * r.dtdyi;
    else
      v = v - (r.p[a + g.sxy] - p0) * r.dt * r.dzc_inv[ks];
  }
  if ((r.fmask >> fc) & 1) v = v + (T)r.force[fc];
  if (ghost) v = fc == 0 ? T(2) * r.bc_u - v : T(2) * T(0) - v;
  return v;
}

// The previous step's projection, forcing and boundary update of u, v, w
// (src/mod_calcVel.cc:253-320, src/mod_updateBound.cc:229-251) in one pass
// over every cell, halos included, into separate arrays (the source's halo
// images must stay unprojected while other threads read them).
template <typename T>
__global__ __launch_bounds__(256) void k_apply_deferred(Geom g, const T* __restrict__ u,
                                                        const T* __restrict__ v,
                                                        const T* __restrict__ w,
                                                        T* __restrict__ du, T* __restrict__ dv,
                                                        T* __restrict__ dw, VelRules<T> r) {
  const long n = (long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % (g.nx + 2)), j = (int)((e / (g.nx + 2)) % (g.ny + 2));
    const int k = (int)(e / ((long)(g.nx + 2) * (g.ny + 2)));
    const long a = g.idx(i, j, k);
    du[a] = rule_load<T>(u, 0, i, j, k, g, r);
    dv[a] = rule_load<T>(v, 1, i, j, k, g, r);
    dw[a] = rule_load<T>(w, 2, i, j, k, g, r);
  }
}

// ---- the closure stage: RK stage (src/mod_calcVel.cc:27-242 in the
// reference's arithmetic) plus the divergence of the eddy-viscosity stress
// 2 nu_t S_ij, with nu_t = nus * softplus(MLP(features)).  k_nut evaluates the
// network once on every cell of the stage's source into nut; the stage kernel
// k_rk_les owns 32 x 8 x kLz cells (kLz = 1: at the coarse grid's 393,216
// cells one-plane blocks give 4x the blocks of four-plane ones and run 7 %
// faster): phase 1 loads nu_t of the 34 x 10 x (kLz + 2) cells
// around them (periodic images in x and y; 0 outside 1 <= k <= nz) into LDS,
// phase 2 updates the velocities.  Staggered stress
// divergence: normal stresses at cell centres, shear stresses at cell edges
// with nu_t averaged from the four cells around the edge; the wall edges carry
// no eddy viscosity (the wall stress is the molecular one).
// ---- register-resident inference (the closure stage) ----------------------
// The layers are computed transposed, hidden x samples: h1^T = W1^T X^T and
// h2^T = W2^T h1^T, a 32-sample tile per wave.  The accumulator of a
// v_mfma_f32_32x32x16_bf16 leaves lane l (half h = l/32) with sample l%32 and
// the 16 hidden units u(j, h) = 8(j/4) + 4h + j%4 (+32 per row block); the B
// operand of the next product wants, in lane l, sample l%32 and 8 consecutive
// k-slots of half h.  Numbering layer 2's k-slots (kb, h, e) as the units
// u(8(kb%2) + e, h) + 32(kb/2) makes every lane's own results its next
// operand: no LDS and no barrier between the layers.  The biases start the
// accumulators; the output dot is split over the halves and joined with one
// lane exchange.  Same rounding points as forward_tile (BF16 features,
// weights and activations; FP32 sums); only the summation order differs.
//
// Input k-slots: half 0 holds features 0..4 (du/dx, du/dy, du/dz, dv/dx,
// dv/dy) in slots 0..4, half 1 features 5..8 (dv/dz, dw/dx, dw/dy, dw/dz) in
// slots 8..11 and the wall distance in slot 12; the other slots are zero.
__device__ __forceinline__ int hid_unit(int j, int h) { return 8 * (j / 4) + 4 * h + (j % 4); }
__device__ __forceinline__ int slot_input(int s) {  // feature of input k-slot s, -1: zero
  return s < 5 ? s : (s < 8 ? -1 : (s < 13 ? s - 3 : -1));
}

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Two FP32 values rounded to BF16 (nearest, ties to even) in one dword, lo | hi << 16.
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}
// BF16 of relu(lo), relu(hi), packed: a negative value rounds to a BF16 with
// its sign bit set, which a 16-bit signed max against 0 clears to +0, so this
// is the BF16 of max(x, 0) in two instructions for the pair.
__device__ __forceinline__ uint32_t pack_relu_bf16(float lo, float hi) {
  const s16x2 q = __builtin_bit_cast(s16x2, pack_bf16(lo, hi));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(q, (s16x2){0, 0}));
}

struct RegMlp {
  bf16x8 a1[2];      // W1^T rows (hidden units l%32 + 32t), k-slots 8h..8h+7
  bf16x8 a2[2][4];   // W2^T rows (units l%32 + 32t), k-slots of block kb
  uint32_t w3[2][8];    // BF16 W3 of the lane's layer-2 rows j = 2m, 2m+1, packed
  float b3;
};

// The hidden biases b1 (lb[0..63]) and b2 (lb[64..127]) are read from LDS into
// the accumulator rows of lane half h, row block t, when a tile starts: the
// rows j = 4q .. 4q+3 are the consecutive units 8q + 4h + 32t .. +3.
__device__ __forceinline__ f32x16 bias_rows(const float* lb, int h, int t) {
  f32x16 c;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 v = *reinterpret_cast<const float4*>(lb + 8 * q + 4 * h + 32 * t);
    c[4 * q] = v.x;
    c[4 * q + 1] = v.y;
    c[4 * q + 2] = v.z;
    c[4 * q + 3] = v.w;
  }
  return c;
}

__device__ inline void load_reg_mlp(const Mlp& w, int lane, RegMlp& r) {
  const int h = lane / 32, m = lane % 32;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int u = m + 32 * t;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int in = slot_input(8 * h + e);
      r.a1[t][e] = in < 0 ? (short)0 : (short)w.w1[in * kH + u];
    }
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int unit = hid_unit(8 * (kb % 2) + e, h) + 32 * (kb / 2);
        r.a2[t][kb][e] = (short)w.w2[unit * kH + u];
      }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      r.w3[t][q] = (uint32_t)w.w3[hid_unit(2 * q, h) + 32 * t] |
                   ((uint32_t)w.w3[hid_unit(2 * q + 1, h) + 32 * t] << 16);
  }
  r.b3 = w.p[oB3];
}

// o of sample lane%32 (in both halves) from the lane's 8 input slots x; lb
// holds the hidden biases (bias_rows).  The output dot runs on
// v_dot2c_f32_bf16 (BF16 h2 times BF16 W3, FP32 sum).
__device__ inline float reg_forward(const RegMlp& r, const float* lb, int h, const bf16x8 b) {
  f32x16 z1[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
    z1[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(r.a1[t], b, bias_rows(lb, h, t), 0, 0, 0);
  bf16x8 b2[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const f32x16& z = z1[kb / 2];
    const int o = 8 * (kb % 2);
    b2[kb] = __builtin_bit_cast(
        bf16x8, (u32x4){pack_relu_bf16(z[o], z[o + 1]), pack_relu_bf16(z[o + 2], z[o + 3]),
                        pack_relu_bf16(z[o + 4], z[o + 5]), pack_relu_bf16(z[o + 6], z[o + 7])});
  }
  // the two row blocks' products alternate, so consecutive MFMAs are independent
  f32x16 z2[2] = {bias_rows(lb + kH, h, 0), bias_rows(lb + kH, h, 1)};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int t = 0; t < 2; ++t)
      z2[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(r.a2[t][kb], b2[kb], z2[t], 0, 0, 0);
  float part = 0.f;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int q = 0; q < 8; ++q)
      part = __builtin_amdgcn_fdot2_f32_bf16(
          __builtin_bit_cast(bf16x2, pack_relu_bf16(z2[t][2 * q], z2[t][2 * q + 1])),
          __builtin_bit_cast(bf16x2, r.w3[t][q]), part, false);
  return part + __shfl_xor(part, 32) + r.b3;
}

constexpr int kLx = 32, kLy = 8, kLz = 1;
constexpr int kEx = kLx + 2, kEy = kLy + 2, kEz = kLz + 2, kExt = kEx * kEy * kEz;
static_assert(kLx * kLy == 64 * kWaves, "one thread per (i, j) column of the tile");

// The 16 field values one half of a sample's inputs reads, as (field, offset
// from the cell a = idx(i,j,k)).  Both halves take three 4-point central
// differences ((p0 + p1) - (p2 + p3), Q0..Q2) and two face differences
// (p0 - p1, D0 and D1), in cell_gradient's operand order:
//   half 0: Q0 du/dy, Q1 du/dz, Q2 dv/dx, D0 du/dx, D1 dv/dy
//   half 1: Q0 dv/dz, Q1 dw/dx, Q2 dw/dy, D0 dw/dz (D1 unused, reads w[a], w[a-sz])
// so the two halves of a wave run the same instructions on different addresses.
template <typename T>
__device__ inline void les_half_loads(const T* u, const T* v, const T* w, long sx, long sz,
                                      int half, const T* fp[16], long fo[16]) {
  const long o0[16] = {sx, sx - 1, -sx, -sx - 1, sz, sz - 1, -sz, -sz - 1,
                       1,  1 - sx, -1,  -1 - sx, 0,  -1,     0,   -sx};
  const long o1[16] = {sz, sz - sx, -sz, -sz - sx, 1,  1 - sz, -1,  -1 - sz,
                       sx, sx - sz, -sx, -sx - sz, 0,  -sz,    0,   -sz};
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const T* f0 = (s < 8 || s == 12 || s == 13) ? u : v;
    const T* f1 = s < 4 ? v : w;
    fp[s] = half ? f1 : f0;
    fo[s] = half ? o1[s] : o0[s];
  }
}

// The values of one field that the closure stage reads around a cell: slot 0
// the cell, 1..6 its face neighbours (+x, -x, +y, -y, +z, -z), 7 and 8 two
// diagonal neighbours, which differ by field (u: (-1,1,0), (-1,0,1); v:
// (1,-1,0), (0,-1,1); w: (1,0,-1), (0,1,-1)).  Every offset tendency() and the
// stress divergence use is one of these.
constexpr int kStencil[3][9][3] = {
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}, {-1, 1, 0}, {-1, 0, 1}},
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}, {1, -1, 0}, {0, -1, 1}},
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}, {1, 0, -1}, {0, 1, -1}}};

__host__ __device__ constexpr int stencil_slot(int f, int di, int dj, int dk) {
  return (di == 0 && dj == 0 && dk == 0)   ? 0
         : (dj == 0 && dk == 0)            ? (di == 1 ? 1 : 2)
         : (di == 0 && dk == 0)            ? (dj == 1 ? 3 : 4)
         : (di == 0 && dj == 0)            ? (dk == 1 ? 5 : 6)
         : (f == 2 ? (dj == 0 ? 7 : 8) : (dk == 0 ? 7 : 8));
}
static_assert(stencil_slot(0, -1, 1, 0) == 7 && stencil_slot(0, -1, 0, 1) == 8 &&
                  stencil_slot(1, 1, -1, 0) == 7 && stencil_slot(1, 0, -1, 1) == 8 &&
                  stencil_slot(2, 1, 0, -1) == 7 && stencil_slot(2, 0, 1, -1) == 8 &&
                  stencil_slot(1, 0, 0, -1) == 6 && stencil_slot(2, -1, 0, 0) == 2,
              "stencil slots follow kStencil");

// One plane's stencil values of the three fields (and, for stage 2, the old
// values of the output arrays at the cell), held in registers.
template <typename T>
struct LesPlane {
  T f[3][9];
  T o[3];
};

template <typename T, bool OLD>
__device__ __forceinline__ void load_les_plane(LesPlane<T>& p, const T* u, const T* v, const T* w,
                                               const T* ou, const T* ov, const T* ow, long a,
                                               long sx, long sz) {
  const T* src[3] = {u, v, w};
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int s = 0; s < 9; ++s)
      p.f[f][s] = src[f][a + kStencil[f][s][0] + kStencil[f][s][1] * sx + kStencil[f][s][2] * sz];
  if (OLD) {
    p.o[0] = ou[a];
    p.o[1] = ov[a];
    p.o[2] = ow[a];
  }
}

// tendency()'s stencil accessor over a LesPlane field (f = 0, 1, 2 for u, v, w).
template <typename T>
struct RegAcc {
  const T* r;
  int f;
  __device__ __forceinline__ T operator()(int di, int dj, int dk) const {
    return r[stencil_slot(f, di, dj, dk)];
  }
};

// One tile's inputs for a lane: the fetched field values, the plane
// coefficients of its cell, and whether the sample is a cell of the domain.
template <typename T>
struct LesIn {
  T v[16];
  T hz, dfi, zs, yw;
  bool valid;
};

// The eddy viscosity of every cell of a stage's source field, in chunks of
// 64*kWaves cells per block (grid-stride): first every thread forms the 10
// features of one cell (all its loads in flight at once) and stores them as
// the BF16 input k-slots of the network in LDS, then every wave runs the
// network in registers (reg_forward: the weights loaded once per block) on its
// 32-cell tiles of the chunk.  The velocity differences and features are formed
// in FP32 (they are rounded to BF16 next), in cell_gradient's operand order and
// features()' scaling and reflection signs (du/dz, dv/dz, dw/dx, dw/dy).  nut
// is interior, x fastest.
constexpr int kNutChunk = 64 * kWaves;
template <typename T>
__global__ __launch_bounds__(64 * kWaves) void k_nut(Geom g, const T* __restrict__ u,
                                                    const T* __restrict__ v,
                                                    const T* __restrict__ w, MeshT<T> m,
                                                    FeatMesh<T> fm, float dxi, float dyi,
                                                    float qdxi, float qdyi, Mlp net, float gs,
                                                    float ys, float nus, float* __restrict__ nut,
                                                    long ncell) {
  __shared__ __align__(16) float lb[2 * kH];            // b1, b2
  __shared__ u32x4 xs[2][kNutChunk];                    // k-slots 0..7 and 8..15 per cell
  for (int q = threadIdx.x; q < 2 * kH; q += kNutChunk) lb[q] = net.p[(q < kH ? oB1 : oB2 - kH) + q];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64, half = lane / 32;
  RegMlp r;
  load_reg_mlp(net, lane, r);
  const long nxy = (long)g.nx * g.ny, sx = g.sx, sz = g.sxy;
  const long nchunk = (ncell + kNutChunk - 1) / kNutChunk;
  // the 30 field values and 4 plane coefficients of this thread's cell of a
  // chunk, loaded one chunk ahead of the features that use them
  const long off[30] = {sx, sx - 1, -sx, -sx - 1, sz, sz - 1, -sz, -sz - 1, 0, -1,           // u
                        1, 1 - sx, -1, -1 - sx, sz, sz - sx, -sz, -sz - sx, 0, -sx,         // v
                        1, 1 - sz, -1, -1 - sz, sx, sx - sz, -sx, -sx - sz, 0, -sz};        // w
  T raw[30];
  T coef[4];
  long ecur = 0;
  auto load = [&](long ch) {
    const long e0 = ch * kNutChunk + threadIdx.x;
    ecur = e0;
    const long e = e0 < ncell ? e0 : ncell - 1;
    const int i = (int)(e % g.nx) + 1, j = (int)((e / g.nx) % g.ny) + 1, k = (int)(e / nxy) + 1;
    const long a = g.idx(i, j, k);
#pragma unroll
    for (int q = 0; q < 30; ++q) raw[q] = (q < 10 ? u : (q < 20 ? v : w))[a + off[q]];
    coef[0] = fm.hz[k];
    coef[1] = m.dzf_inv[k];
    coef[2] = fm.zs[k];
    coef[3] = fm.yw[k];
  };
  long ch = blockIdx.x;
  if (ch < nchunk) load(ch);
  for (; ch < nchunk; ch += gridDim.x) {
    {
      float x[30];
#pragma unroll
      for (int q = 0; q < 30; ++q) x[q] = (float)raw[q];
      const float hz = (float)coef[0], dfi = (float)coef[1], zs = (float)coef[2], yw = (float)coef[3];
      const float qu_y = (x[0] + x[1]) - (x[2] + x[3]), qu_z = (x[4] + x[5]) - (x[6] + x[7]);
      const float qv_x = (x[10] + x[11]) - (x[12] + x[13]), qv_z = (x[14] + x[15]) - (x[16] + x[17]);
      const float qw_x = (x[20] + x[21]) - (x[22] + x[23]), qw_y = (x[24] + x[25]) - (x[26] + x[27]);
      const float du_x = x[8] - x[9], dv_y = x[18] - x[19], dw_z = x[28] - x[29];
      const bool valid = ecur < ncell;
      const float f0 = valid ? (du_x * dxi) * gs : 0.f;          // du/dx
      const float f1 = valid ? (qu_y * qdyi) * gs : 0.f;         // du/dy
      const float f2 = valid ? ((qu_z * hz) * gs) * zs : 0.f;    // du/dz
      const float f3 = valid ? (qv_x * qdxi) * gs : 0.f;         // dv/dx
      const float f4 = valid ? (dv_y * dyi) * gs : 0.f;          // dv/dy
      const float f5 = valid ? ((qv_z * hz) * gs) * zs : 0.f;    // dv/dz
      const float f6 = valid ? ((qw_x * qdxi) * gs) * zs : 0.f;  // dw/dx
      const float f7 = valid ? ((qw_y * qdyi) * gs) * zs : 0.f;  // dw/dy
      const float f8 = valid ? (dw_z * dfi) * gs : 0.f;          // dw/dz
      const float f9 = valid ? yw * ys : 0.f;                    // wall distance
      xs[0][threadIdx.x] = (u32x4){pack_bf16(f0, f1), pack_bf16(f2, f3), pack_bf16(f4, 0.f), 0u};
      xs[1][threadIdx.x] = (u32x4){pack_bf16(f5, f6), pack_bf16(f7, f8), pack_bf16(f9, 0.f), 0u};
    }
    __syncthreads();  // also orders lb before the first tile
    if (ch + gridDim.x < nchunk) load(ch + gridDim.x);
#pragma unroll
    for (int q = 0; q < kNutChunk / 32 / kWaves; ++q) {
      const int t = wave + q * kWaves;  // tile of the chunk
      const bf16x8 b = __builtin_bit_cast(bf16x8, xs[half][t * 32 + lane % 32]);
      const float o = reg_forward(r, lb, half, b);
      const long e = ch * kNutChunk + t * 32 + lane % 32;
      if (lane < 32 && e < ncell) nut[e] = softplus(o) * nus;
    }
    __syncthreads();  // xs is rewritten by the next chunk
  }
}

// u1's boundary update (src/mod_updateBound.cc:229-251, k_bc_vel's closed
// form) written by the interior cells it copies: every x/y halo image of the
// cell on its plane, the ghost planes 0 and nz+1 by the wall rules, and w = 0
// on the wall plane nz.  Each halo cell has one interior source, so exactly one
// thread writes it.
template <typename T>
__device__ __forceinline__ void store_with_halos(const Geom& g, T* du, T* dv, T* dw, int i,
                                                 int j, int k, T a, T b, T c, T bc_u) {
  const T zero = T(0), two = T(2);
  if (k == g.nz) c = zero;
  const int xi[3] = {i, i == 1 ? g.nx + 1 : -1, i == g.nx ? 0 : -1};
  const int yj[3] = {j, j == 1 ? g.ny + 1 : -1, j == g.ny ? 0 : -1};
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (xi[p] < 0 || yj[q] < 0) continue;
      const long d = g.idx(xi[p], yj[q], k);
      du[d] = a;
      dv[d] = b;
      dw[d] = c;
      if (k == 1) {
        const long e = g.idx(xi[p], yj[q], 0);
        du[e] = two * bc_u - a;
        dv[e] = two * zero - b;
        dw[e] = zero;
      }
      if (k == g.nz) {
        const long e = g.idx(xi[p], yj[q], g.nz + 1);
        du[e] = two * bc_u - a;
        dv[e] = two * zero - b;
        dw[e] = zero;
      }
    }
}

// With `halo` stage 1 also writes u1's boundary (store_with_halos).  With
// `acc` stage 2 sums its new u times dzflzi(k) in FP64 into partial[block]
// (the bulk mean of the next projected velocity, as k_rk's).
template <typename T, int STAGE>
__global__ __launch_bounds__(64 * kWaves) void k_rk_les(
    Geom g, const T* __restrict__ su, const T* __restrict__ sv, const T* __restrict__ sw,
    T* __restrict__ du, T* __restrict__ dv, T* __restrict__ dw, MeshT<T> m, T dxi, T dyi,
    T rei, T dt, T ucrf, const float* __restrict__ nut, int ntx, int nty, int halo, T bc_u,
    int acc, const double* __restrict__ dzflzi, double* __restrict__ partial) {
  __shared__ __align__(16) float nu[kExt];  // 16-byte aligned: reused as FP64 sums
  const int b = blockIdx.x;
  const int ti = b % ntx, tj = (b / ntx) % nty, tk = b / (ntx * nty);
  const int i0 = ti * kLx, j0 = tj * kLy, k0 = tk * kLz;  // extended cell (0,0,0) = (i0,j0,k0)
  // phase 1: the eddy viscosity (k_nut) of the extended block into LDS: x/y
  // halo cells from their periodic images, 0 outside the walls
  for (int e = threadIdx.x; e < kExt; e += 64 * kWaves) {
    const int ex = e % kEx, ey = (e / kEx) % kEy, ez = e / (kEx * kEy);
    int i = i0 + ex, j = j0 + ey;
    const int k = k0 + ez;
    const bool valid = i <= g.nx + 1 && j <= g.ny + 1 && k >= 1 && k <= g.nz;
    i = i == 0 ? g.nx : (i > g.nx ? 1 : i);
    j = j == 0 ? g.ny : (j > g.ny ? 1 : j);
    nu[e] = valid ? nut[((long)(k - 1) * g.ny + (j - 1)) * g.nx + (i - 1)] : 0.f;
  }
  __syncthreads();
  // phase 2: the stage update of the block's cells, a column of kLz planes
  // per thread; the stencil values of the next plane are loaded before the
  // current plane is computed.  Planes beyond nz (a ragged last block) are
  // computed on plane nz's values and not stored.
  const int tx = threadIdx.x % kLx, ty = threadIdx.x / kLx;
  const bool inside = i0 + tx + 1 <= g.nx && j0 + ty + 1 <= g.ny;
  // threads past the grid work on its last cell and store nothing
  const int i = inside ? i0 + tx + 1 : g.nx, j = inside ? j0 + ty + 1 : g.ny;
  const T two = T(2), qtr = T(0.25);
  double msum = 0.0;
  auto plane_of = [&](int kz) { const int k = k0 + kz + 1; return k < g.nz ? k : g.nz; };
  LesPlane<T> cur, nxt;
  load_les_plane<T, STAGE == 2>(cur, su, sv, sw, du, dv, dw, g.idx(i, j, plane_of(0)), g.sx, g.sxy);
#pragma unroll
  for (int kz = 0; kz < kLz; ++kz) {
    const int k = plane_of(kz);
    const bool live = inside && k0 + kz + 1 <= g.nz;
    if (kz + 1 < kLz)
      load_les_plane<T, STAGE == 2>(nxt, su, sv, sw, du, dv, dw, g.idx(i, j, plane_of(kz + 1)),
                                    g.sx, g.sxy);
    const long a = g.idx(i, j, k);
    const RegAcc<T> U{cur.f[0], 0}, V{cur.f[1], 1}, W{cur.f[2], 2};
    T tu, tv, tw;
    tendency<T>(U, V, W, m, k, dxi, dyi, rei, ucrf, tu, tv, tw);
    const int c = (tx + 1) + kEx * ((ty + 1) + kEy * (kz + 1));
    auto N = [&](int di, int dj, int dk) -> T { return (T)nu[c + di + kEx * (dj + kEy * dk)]; };
    const T dci = m.dzc_inv[k], dcm = m.dzc_inv[k - 1], dfi = m.dzf_inv[k];
    const bool top = k == g.nz, bot = k == 1;
    T s_u, s_v, s_w;
    {  // x-momentum at the u face (i+1/2, j, k)
      const T sE = (U(1, 0, 0) - U(0, 0, 0)) * dxi, sW = (U(0, 0, 0) - U(-1, 0, 0)) * dxi;
      const T fxx = ((N(1, 0, 0) * sE - N(0, 0, 0) * sW) * two) * dxi;
      const T nN = qtr * (((N(0, 0, 0) + N(1, 0, 0)) + N(0, 1, 0)) + N(1, 1, 0));
      const T nS = qtr * (((N(0, -1, 0) + N(1, -1, 0)) + N(0, 0, 0)) + N(1, 0, 0));
      const T gN = (U(0, 1, 0) - U(0, 0, 0)) * dyi + (V(1, 0, 0) - V(0, 0, 0)) * dxi;
      const T gS = (U(0, 0, 0) - U(0, -1, 0)) * dyi + (V(1, -1, 0) - V(0, -1, 0)) * dxi;
      const T fxy = (nN * gN - nS * gS) * dyi;
      // the wall edges carry no eddy viscosity (all neighbours are in the
      // block's LDS, so both sides are computed and one is selected)
      const T nTf = qtr * (((N(0, 0, 0) + N(1, 0, 0)) + N(0, 0, 1)) + N(1, 0, 1));
      const T nBf = qtr * (((N(0, 0, -1) + N(1, 0, -1)) + N(0, 0, 0)) + N(1, 0, 0));
      const T nT = top ? T(0) : nTf, nB = bot ? T(0) : nBf;
      const T gT = (U(0, 0, 1) - U(0, 0, 0)) * dci + (W(1, 0, 0) - W(0, 0, 0)) * dxi;
      const T gB = (U(0, 0, 0) - U(0, 0, -1)) * dcm + (W(1, 0, -1) - W(0, 0, -1)) * dxi;
      const T fxz = (nT * gT - nB * gB) * dfi;
      s_u = (fxx + fxy) + fxz;
    }
    {  // y-momentum at the v face (i, j+1/2, k)
      const T nE = qtr * (((N(0, 0, 0) + N(1, 0, 0)) + N(0, 1, 0)) + N(1, 1, 0));
      const T nW = qtr * (((N(-1, 0, 0) + N(0, 0, 0)) + N(-1, 1, 0)) + N(0, 1, 0));
      const T gE = (U(0, 1, 0) - U(0, 0, 0)) * dyi + (V(1, 0, 0) - V(0, 0, 0)) * dxi;
      const T gW = (U(-1, 1, 0) - U(-1, 0, 0)) * dyi + (V(0, 0, 0) - V(-1, 0, 0)) * dxi;
      const T fyx = (nE * gE - nW * gW) * dxi;
      const T sN = (V(0, 1, 0) - V(0, 0, 0)) * dyi, sS = (V(0, 0, 0) - V(0, -1, 0)) * dyi;
      const T fyy = ((N(0, 1, 0) * sN - N(0, 0, 0) * sS) * two) * dyi;
      const T nTf = qtr * (((N(0, 0, 0) + N(0, 1, 0)) + N(0, 0, 1)) + N(0, 1, 1));
      const T nBf = qtr * (((N(0, 0, -1) + N(0, 1, -1)) + N(0, 0, 0)) + N(0, 1, 0));
      const T nT = top ? T(0) : nTf, nB = bot ? T(0) : nBf;
      const T gT = (V(0, 0, 1) - V(0, 0, 0)) * dci + (W(0, 1, 0) - W(0, 0, 0)) * dyi;
      const T gB = (V(0, 0, 0) - V(0, 0, -1)) * dcm + (W(0, 1, -1) - W(0, 0, -1)) * dyi;
      const T fyz = (nT * gT - nB * gB) * dfi;
      s_v = (fyx + fyy) + fyz;
    }
    {  // z-momentum at the w face (i, j, k+1/2); w(nz) is the wall value
      const T nE = qtr * (((N(0, 0, 0) + N(1, 0, 0)) + N(0, 0, 1)) + N(1, 0, 1));
      const T nW = qtr * (((N(-1, 0, 0) + N(0, 0, 0)) + N(-1, 0, 1)) + N(0, 0, 1));
      const T gE = (U(0, 0, 1) - U(0, 0, 0)) * dci + (W(1, 0, 0) - W(0, 0, 0)) * dxi;
      const T gW = (U(-1, 0, 1) - U(-1, 0, 0)) * dci + (W(0, 0, 0) - W(-1, 0, 0)) * dxi;
      const T fzx = (nE * gE - nW * gW) * dxi;
      const T nN = qtr * (((N(0, 0, 0) + N(0, 1, 0)) + N(0, 0, 1)) + N(0, 1, 1));
      const T nS = qtr * (((N(0, -1, 0) + N(0, 0, 0)) + N(0, -1, 1)) + N(0, 0, 1));
      const T gN = (V(0, 0, 1) - V(0, 0, 0)) * dci + (W(0, 1, 0) - W(0, 0, 0)) * dyi;
      const T gS = (V(0, -1, 1) - V(0, -1, 0)) * dci + (W(0, 0, 0) - W(0, -1, 0)) * dyi;
      const T fzy = (nN * gN - nS * gS) * dyi;
      const T sT = (W(0, 0, 1) - W(0, 0, 0)) * m.dzf_inv[k + 1];
      const T sB = (W(0, 0, 0) - W(0, 0, -1)) * dfi;
      const T fzz = ((N(0, 0, 1) * sT - N(0, 0, 0) * sB) * two) * dci;
      s_w = top ? T(0) : (fzx + fzy) + fzz;
    }
    if (live) {
      if (STAGE == 1) {
        const T n1 = U(0, 0, 0) + dt * tu + dt * s_u;
        const T n2 = V(0, 0, 0) + dt * tv + dt * s_v;
        const T n3 = W(0, 0, 0) + dt * tw + dt * s_w;
        if (halo) {
          store_with_halos(g, du, dv, dw, i, j, k, n1, n2, n3, bc_u);
        } else {
          du[a] = n1;
          dv[a] = n2;
          dw[a] = n3;
        }
      } else {
        const T nu = (U(0, 0, 0) + dt * tu + cur.o[0] + dt * s_u) * T(0.5);
        du[a] = nu;
        dv[a] = (V(0, 0, 0) + dt * tv + cur.o[1] + dt * s_v) * T(0.5);
        dw[a] = (W(0, 0, 0) + dt * tw + cur.o[2] + dt * s_w) * T(0.5);
        if (acc) msum += (double)nu * dzflzi[k];
      }
    }
    if (kz + 1 < kLz) cur = nxt;
  }
  if (STAGE == 2 && acc) {
    // nu (the eddy viscosity block) is free: every thread passed its reads
    __syncthreads();
    double* red = reinterpret_cast<double*>(nu);
    static_assert(sizeof(nu) >= 64 * kWaves * sizeof(double), "reduction scratch");
    red[threadIdx.x] = msum;
    __syncthreads();
    for (int h = 32 * kWaves; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[b] = red[0];
  }
}

}  // namespace
}  // namespace pll

using namespace pll;

struct pll_closure {
  float* p = nullptr;   // FP32 master parameters
  float* m1 = nullptr;  // Adam moments
  float* m2 = nullptr;
  uint16_t* wb = nullptr;  // BF16 W1 | W2 | W3
  float* partial = nullptr;
  long npartial_blocks = 0;  // partial gradients, one per k_mlp block
  float* grad = nullptr;    // kNParam + 1 (loss sum last)
  float* chunks = nullptr;  // first-level gradient sums, (kMaxWaves / kRedChunk) x (kNParam + 1)
  long step = 0;
  hipStream_t stream = nullptr;
  int device = 0;              // the device current at creation (that of stream)
  hipEvent_t ready = nullptr;  // orders a context's closure stages after training
  long version = 0;            // bumped by every asynchronous change of the weights
  // feature and output scales: inputs g*gs and (wall distance)*ys, nu_t = nus*softplus(o)
  double gs = 1.0, ys = 1.0, nus = 1.0;
};

namespace {

Mlp weights(const pll_closure* c) {
  return Mlp{c->p, c->wb + kWb1, c->wb + kWb2, c->wb + kWb3, c->wb + kWbB1, c->wb + kWbB2};
}

int ensure_partial(pll_closure* c, long nblocks) {
  if (nblocks <= c->npartial_blocks) return PLL_OK;
  if (c->partial) (void)hipFree(c->partial);
  c->partial = nullptr;
  PLL_HIP_TRY(hipMalloc(&c->partial, (size_t)nblocks * (kNParam + 1) * sizeof(float)));
  c->npartial_blocks = nblocks;
  return PLL_OK;
}

// Waves of a k_mlp grid for n samples (a multiple of kWaves, at most
// kMaxWaves) and the 32-sample tiles each wave runs (wave q takes tiles q,
// q + nwaves, ...).
long waves_for(long n) {
  const long tiles = (n + 31) / 32;
  const long waves = std::min(tiles, kMaxWaves);
  return (waves + kWaves - 1) / kWaves * kWaves;
}
int tiles_per_wave(long n, long nwaves) {
  const long tiles = (n + 31) / 32;
  return (int)((tiles + nwaves - 1) / nwaves);
}

// The gradient into chunk sums; with `final` also summed into c->grad (the
// training step leaves the last level to k_adam).  Returns the chunk count.
long gradient_chunks(pll_closure* c, const float* x, const float* y, long n, bool final, int* rc) {
  const long nw = waves_for(n);
  *rc = ensure_partial(c, nw / kWaves);
  if (*rc) return 0;
  hipLaunchKernelGGL((k_mlp<true>), dim3((unsigned)(nw / kWaves)), dim3(64 * kWaves), 0, c->stream,
                     weights(c), x, y, n, static_cast<float*>(nullptr), c->partial,
                     tiles_per_wave(n, nw));
  const long nblk = nw / kWaves;  // one partial per block
  const long nchunk = (nblk + kRedChunk - 1) / kRedChunk;
  hipLaunchKernelGGL(k_grad_reduce, dim3((kNParam + 1 + 255) / 256, (unsigned)nchunk), dim3(256), 0,
                     c->stream, c->partial, nblk, c->chunks);
  if (final)
    hipLaunchKernelGGL(k_grad_final, dim3((kNParam + 1 + 255) / 256), dim3(256), 0, c->stream,
                       c->chunks, nchunk, c->grad);
  const hipError_t e = hipGetLastError();
  *rc = e == hipSuccess ? PLL_OK : hip_fail(e, "closure gradient kernels");
  return nchunk;
}

int gradient(pll_closure* c, const float* x, const float* y, long n) {
  int rc = PLL_OK;
  gradient_chunks(c, x, y, n, true, &rc);
  return rc;
}

// ---- allocation-free gradient ------------------------------------------------
// The buffers of pll_closure_gradient_noalloc live in one file-scope device
// array, one copy per GPU for the life of the process: the per-block partial
// gradients of the largest training grid (kMaxWaves / kWaves blocks), the
// first-level chunk sums and the summed gradient with the loss, each
// kNParam + 1 floats -- 6.1 MB.  Every call on a GPU shares that GPU's copy;
// each call writes every element it reads before reading it, so nothing has
// to start at zero, and calls that overlap on one GPU race.
constexpr long kScratchPartials = kMaxWaves / kWaves;
constexpr long kScratchChunks = kMaxWaves / kRedChunk;
constexpr long kScratchFloats = (kScratchPartials + kScratchChunks + 1) * (kNParam + 1);
static __device__ __align__(256) float scratch[kScratchFloats];

// Device address of `scratch` on each GPU, resolved once per device with that
// device current (the host-side name is only a stand-in for the symbol).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

int scratch_on(int dev, float** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(PLL_ERR_STATE, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return PLL_OK;
}

// Label regularisation: eps of nu* in units of the squared scaled strain rate,
// and the largest label.
constexpr double kLabelEps = 1e-4;
constexpr double kLabelMax = 50.0;

template <typename T>
int feat_setup(pll_context* c) {
  if (c->d_feat) return PLL_OK;
  const int nz = c->g.nz;
  const double lz = c->cfg.lz;
  std::vector<T> h(3 * (nz + 2), T(0));
  for (int k = 1; k <= nz; ++k) {
    const double zc = c->zc_global[k];
    h[k] = (T)(0.5 / (c->dzc[k] + c->dzc[k - 1]));
    h[(nz + 2) + k] = (T)std::min(zc, lz - zc);
    h[2 * (nz + 2) + k] = (T)(zc < 0.5 * lz ? 1.0 : -1.0);
  }
  PLL_HIP_TRY(ctx_malloc(c, &c->d_feat, h.size() * sizeof(T)));
  PLL_HIP_TRY(hipMemcpy(c->d_feat, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return PLL_OK;
}
int feat_setup_any(pll_context* c) {
  return c->prec == PLL_FP64 ? feat_setup<double>(c) : feat_setup<float>(c);
}

template <typename T>
FeatMesh<T> feat_mesh(const pll_context* c) {
  const T* d = static_cast<const T*>(c->d_feat);
  const int n = c->g.nz + 2;
  return FeatMesh<T>{d, d + n, d + 2 * n};
}

template <typename T>
MeshT<T> mesh_t(const pll_context* c) {
  const T* d = static_cast<const T*>(c->d_mesh);
  const int n = c->g.nz + 2;
  return MeshT<T>{d, d + n, d + 2 * n, d + 3 * n};
}

template <typename T>
T* fld(const pll_context* c, int f) {
  return static_cast<T*>(c->fld[f]);
}

// The closure stage.  Stage 1 with fold (and corr) first applies the previous
// step's forcing (and projection) and boundary update in one pass
// (k_apply_deferred), and with bcload writes u1's boundary update itself;
// stage 2 with acc sums the bulk-mean partials of its new u for the next
// step's forcing.  The sources are read as stored, halos included.
template <typename T>
int rk_closure_t(pll_context* c, int stage, int fold, int corr, int acc, int bcload) {
  pll_closure* cl = c->closure;
  if (c->closure_version != cl->version) {
    // the weights changed on the closure's stream (training) since this
    // context last waited for them
    PLL_HIP_TRY(hipEventRecord(cl->ready, cl->stream));
    PLL_HIP_TRY(hipStreamWaitEvent(c->stream, cl->ready, 0));
    c->closure_version = cl->version;
  }
  const Geom g = c->g;
  const int ntx = (g.nx + kLx - 1) / kLx, nty = (g.ny + kLy - 1) / kLy, ntz = (g.nz + kLz - 1) / kLz;
  const dim3 grid((unsigned)ntx * nty * ntz), block(64 * kWaves);
  const T dxi = (T)c->dx_inv, dyi = (T)c->dy_inv;
  const T rei = (T)c->re_inv, dt = (T)c->dt, ucrf = (T)c->u_crf;
  const float nus = (float)cl->nus;
  const T* su = fld<T>(c, stage == 1 ? PLL_U : PLL_U1);
  const T* sv = fld<T>(c, stage == 1 ? PLL_V : PLL_V1);
  const T* sw = fld<T>(c, stage == 1 ? PLL_W : PLL_W1);
  // the network on every cell of the stage's source (one wave per 32-cell
  // tile, up to kMaxWaves waves), then the stage with that eddy viscosity
  const long ncell = c->nspec;
  const long nblk = std::min((ncell + kNutChunk - 1) / kNutChunk, 512L);
  if (acc && (stage != 2 || (long)grid.x > c->npartial))
    return fail(PLL_ERR_STATE, "the bulk-mean partials are stage 2's and must fit the scratch");
  if (corr && !fold) return fail(PLL_ERR_STATE, "deferred projection needs folded loads");
  if (bcload && stage != 1) return fail(PLL_ERR_STATE, "the closure's stage 2 reads an updated u1");
  const T bc_u = (T)(0.0 - c->u_crf);  // src/mod_updateBound.cc:80
  if (stage == 1 && fold) {
    // the previous step's projection, forcing and boundary update into the
    // spare velocity buffers, which then become u, v, w
    if (!c->fld_alt[0] || !c->fld_alt[1] || !c->fld_alt[2])
      return fail(PLL_ERR_STATE, "no spare velocity buffers");
    const int fmask = (c->is_forced[0] ? 1 : 0) | (c->is_forced[1] ? 2 : 0) | (c->is_forced[2] ? 4 : 0);
    const VelRules<T> r{1, corr, fld<T>(c, PLL_P), mesh_t<T>(c).dzc_inv,
                        (T)(c->dt * c->dx_inv), (T)(c->dt * c->dy_inv),  // src/mod_calcVel.cc:256-258
                        dt, c->d_scalar + 3, fmask, bc_u};
    const long n = (long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    TimerScope ts(c, K_CORRECT);
    hipLaunchKernelGGL((k_apply_deferred<T>), dim3((unsigned)std::min((n + 255) / 256, 8192L)),
                       dim3(256), 0, c->stream, g, su, sv, sw, static_cast<T*>(c->fld_alt[0]),
                       static_cast<T*>(c->fld_alt[1]), static_cast<T*>(c->fld_alt[2]), r);
    PLL_HIP_TRY(hipGetLastError());
    for (int d = 0; d < 3; ++d) std::swap(c->fld[PLL_U + d], c->fld_alt[d]);
    su = fld<T>(c, PLL_U);
    sv = fld<T>(c, PLL_V);
    sw = fld<T>(c, PLL_W);
  }
  const double* dzflzi = c->d_mesh64 + 3 * (c->g.nz + 2);
  {
    TimerScope ts(c, K_RK_CLOSURE);
    hipLaunchKernelGGL((k_nut<T>), dim3((unsigned)nblk), block, 0, c->stream, g, su, sv, sw,
                       mesh_t<T>(c), feat_mesh<T>(c), (float)c->dx_inv, (float)c->dy_inv,
                       (float)(0.25 * c->dx_inv), (float)(0.25 * c->dy_inv), weights(cl),
                       (float)cl->gs, (float)cl->ys, nus, c->d_nut, ncell);
    if (stage == 1)
      hipLaunchKernelGGL((k_rk_les<T, 1>), grid, block, 0, c->stream, g, su, sv, sw,
                         fld<T>(c, PLL_U1), fld<T>(c, PLL_V1), fld<T>(c, PLL_W1), mesh_t<T>(c), dxi,
                         dyi, rei, dt, ucrf, c->d_nut, ntx, nty, bcload, bc_u, 0, dzflzi,
                         c->d_partial);
    else
      hipLaunchKernelGGL((k_rk_les<T, 2>), grid, block, 0, c->stream, g, su, sv, sw,
                         fld<T>(c, PLL_U), fld<T>(c, PLL_V), fld<T>(c, PLL_W), mesh_t<T>(c), dxi,
                         dyi, rei, dt, ucrf, c->d_nut, ntx, nty, 0, bc_u, acc, dzflzi,
                         c->d_partial);
  }
  PLL_HIP_TRY(hipGetLastError());
  // forceVel's bulk mean and force of u (src/mod_calcVel.cc:287,297)
  return acc ? launch_force_from_partials(c, (long)grid.x) : PLL_OK;
}

// Blocks of the closure stage kernel (its bulk-mean partials must fit d_partial).
long closure_stage_blocks(const pll_context* c) {
  const Geom& g = c->g;
  return (long)((g.nx + kLx - 1) / kLx) * ((g.ny + kLy - 1) / kLy) * ((g.nz + kLz - 1) / kLz);
}

// The fine and coarse grids nest: integer factors, the same box, and every
// coarse z cell the union of fz fine ones.
int check_nesting(const pll_context* f, const pll_context* c, int* fx, int* fy, int* fz) {
  if (!f || !c) return fail(PLL_ERR_ARG, "null context");
  const pll_config &a = f->cfg, &b = c->cfg;
  if (a.nx % b.nx || a.ny % b.ny || a.nz % b.nz)
    return fail(PLL_ERR_ARG, "fine grid sizes must be integer multiples of the coarse ones");
  auto same = [](double x, double y) { return std::fabs(x - y) <= 1e-12 * std::fabs(y); };
  if (!same(a.lx, b.lx) || !same(a.ly, b.ly) || !same(a.lz, b.lz))
    return fail(PLL_ERR_ARG, "fine and coarse boxes differ");
  *fx = a.nx / b.nx;
  *fy = a.ny / b.ny;
  *fz = a.nz / b.nz;
  for (int K = 1; K <= b.nz; ++K) {
    double s = 0.0;
    for (int q = 1; q <= *fz; ++q) s += f->dzf[(K - 1) * *fz + q];
    if (!(std::fabs(s - c->dzf[K]) <= 1e-10 * c->dzf[K]))
      return fail(PLL_ERR_ARG, "coarse z cells are not unions of fine ones");
  }
  return PLL_OK;
}

template <typename Tf, typename Tc>
int filter_t(pll_context* f, pll_context* c, int fx, int fy, int fz) {
  const long n = (long)c->g.nx * c->g.ny * c->g.nz;
  hipLaunchKernelGGL((k_filter<Tf, Tc>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     f->g, fld<Tf>(f, PLL_U), fld<Tf>(f, PLL_V), fld<Tf>(f, PLL_W), f->d_mesh64,
                     c->g, fld<Tc>(c, PLL_U), fld<Tc>(c, PLL_V), fld<Tc>(c, PLL_W), fx, fy, fz);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename Tf, typename Tc>
int samples_t(pll_context* f, pll_context* c, const pll_closure* cl, int fx, int fy, int fz,
              float* x, float* y) {
  const long n = (long)c->g.nx * c->g.ny * c->g.nz;
  const double eps = kLabelEps / (cl->gs * cl->gs);
  hipLaunchKernelGGL((k_samples<Tf, Tc>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     c->stream, f->g, fld<Tf>(f, PLL_U), fld<Tf>(f, PLL_V), fld<Tf>(f, PLL_W),
                     f->d_mesh64, c->g, fld<Tc>(c, PLL_U), fld<Tc>(c, PLL_V), fld<Tc>(c, PLL_W),
                     mesh_t<Tc>(c).dzf_inv, feat_mesh<Tc>(c), (Tc)c->dx_inv, (Tc)c->dy_inv,
                     (Tc)(0.25 * c->dx_inv), (Tc)(0.25 * c->dy_inv), (Tc)cl->gs, (Tc)cl->ys,
                     cl->nus, eps, kLabelMax, fx, fy, fz, x, y);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

// filter of the fine context's u, v, w into the coarse one, with its halos
int filter(pll_context* f, pll_context* c) {
  int fx, fy, fz;
  int rc = check_nesting(f, c, &fx, &fy, &fz);
  if (rc) return rc;
  PLL_HIP_TRY(hipStreamSynchronize(f->stream));
  if (f->prec == PLL_FP64)
    rc = c->prec == PLL_FP64 ? filter_t<double, double>(f, c, fx, fy, fz)
                             : filter_t<double, float>(f, c, fx, fy, fz);
  else
    rc = c->prec == PLL_FP64 ? filter_t<float, double>(f, c, fx, fy, fz)
                             : filter_t<float, float>(f, c, fx, fy, fz);
  if (rc) return rc;
  return launch_bc_vel(c, 0);
}

}  // namespace

namespace pll {
int launch_rk_closure(pll_context* c, int stage, int fold, int corr, int acc, int bcload) {
  int rc = feat_setup_any(c);
  if (rc) return rc;
  if (!c->d_nut) PLL_HIP_TRY(ctx_malloc(c, &c->d_nut, (size_t)c->nspec * sizeof(float)));
  return c->prec == PLL_FP64 ? rk_closure_t<double>(c, stage, fold, corr, acc, bcload)
                             : rk_closure_t<float>(c, stage, fold, corr, acc, bcload);
}
bool closure_fused_available(const pll_context* c) {
  return closure_stage_blocks(c) <= c->npartial;
}
}  // namespace pll

extern "C" {

int pll_closure_set_scales(pll_closure* c, double gscale, double yscale, double nuscale) {
  if (!c) return fail(PLL_ERR_ARG, "null closure");
  if (!(gscale > 0) || !(yscale > 0) || !(nuscale > 0))
    return fail(PLL_ERR_ARG, "closure scales must be positive");
  c->gs = gscale;
  c->ys = yscale;
  c->nus = nuscale;
  return PLL_OK;
}

int pll_closure_get_scales(pll_closure* c, double* gscale, double* yscale, double* nuscale) {
  if (!c) return fail(PLL_ERR_ARG, "null closure");
  if (gscale) *gscale = c->gs;
  if (yscale) *yscale = c->ys;
  if (nuscale) *nuscale = c->nus;
  return PLL_OK;
}

// File: "PLLCLOS1", int32 parameter count, that many FP32 parameters, then
// the three FP64 scales (gscale, yscale, nuscale); host byte order.
static const char kClosureMagic[8] = {'P', 'L', 'L', 'C', 'L', 'O', 'S', '1'};

int pll_closure_save(pll_closure* c, const char* path) {
  if (!c || !path) return fail(PLL_ERR_ARG, "null argument");
  std::vector<float> p(kNParam);
  int rc = pll_closure_get_params(c, p.data());
  if (rc) return rc;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(PLL_ERR_IO, std::string("cannot write ") + path);
  const int32_t n = kNParam;
  const double sc[3] = {c->gs, c->ys, c->nus};
  const bool ok = std::fwrite(kClosureMagic, 1, 8, f) == 8 && std::fwrite(&n, 4, 1, f) == 1 &&
                  std::fwrite(p.data(), 4, kNParam, f) == (size_t)kNParam &&
                  std::fwrite(sc, 8, 3, f) == 3;
  if (std::fclose(f) != 0 || !ok) return fail(PLL_ERR_IO, std::string("short write to ") + path);
  return PLL_OK;
}

int pll_closure_load(const char* path, pll_closure** out) {
  if (!path || !out) return fail(PLL_ERR_ARG, "null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(PLL_ERR_IO, std::string("cannot read ") + path);
  char magic[8];
  int32_t n = 0;
  std::vector<float> p(kNParam);
  double sc[3];
  const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kClosureMagic, 8) == 0 &&
                  std::fread(&n, 4, 1, f) == 1 && n == kNParam &&
                  std::fread(p.data(), 4, kNParam, f) == (size_t)kNParam &&
                  std::fread(sc, 8, 3, f) == 3;
  std::fclose(f);
  if (!ok) return fail(PLL_ERR_IO, std::string(path) + " is not a closure file of this network");
  pll_closure* c = nullptr;
  int rc = pll_closure_create(0, &c);
  if (!rc) rc = pll_closure_set_params(c, p.data());
  if (!rc) rc = pll_closure_set_scales(c, sc[0], sc[1], sc[2]);
  if (rc) {
    pll_closure_destroy(c);
    return rc;
  }
  *out = c;
  return PLL_OK;
}

int pll_closure_attach(pll_context* ctx, pll_closure* c) {
  if (!ctx) return fail(PLL_ERR_ARG, "null context");
  if (c) {
    int rc = feat_setup_any(ctx);
    if (rc) return rc;
    if (!ctx->d_nut) PLL_HIP_TRY(ctx_malloc(ctx, &ctx->d_nut, (size_t)ctx->nspec * sizeof(float)));
    PLL_HIP_TRY(hipMemsetAsync(ctx->d_nut, 0, (size_t)ctx->nspec * sizeof(float), ctx->stream));
  }
  ctx->closure = c;
  ctx->closure_version = -1;
  return PLL_OK;
}

int pll_closure_nut(pll_context* ctx, double* host) {
  if (!ctx || !host) return fail(PLL_ERR_ARG, "null argument");
  if (!ctx->d_nut) return fail(PLL_ERR_STATE, "no closure attached to this context");
  std::vector<float> h((size_t)ctx->nspec);
  PLL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  PLL_HIP_TRY(hipMemcpy(h.data(), ctx->d_nut, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < h.size(); ++q) host[q] = (double)h[q];
  return PLL_OK;
}

int pll_closure_filter(pll_context* fine, pll_context* coarse) {
  int rc = filter(fine, coarse);
  if (rc) return rc;
  PLL_HIP_TRY(hipStreamSynchronize(coarse->stream));
  return PLL_OK;
}

int pll_closure_samples(pll_context* fine, pll_context* coarse, pll_closure* cl, float* x,
                        float* y) {
  if (!cl || !x || !y) return fail(PLL_ERR_ARG, "null argument");
  int rc = filter(fine, coarse);
  if (rc) return rc;
  if ((rc = feat_setup_any(coarse))) return rc;
  int fx, fy, fz;
  check_nesting(fine, coarse, &fx, &fy, &fz);
  if (fine->prec == PLL_FP64)
    rc = coarse->prec == PLL_FP64 ? samples_t<double, double>(fine, coarse, cl, fx, fy, fz, x, y)
                                  : samples_t<double, float>(fine, coarse, cl, fx, fy, fz, x, y);
  else
    rc = coarse->prec == PLL_FP64 ? samples_t<float, double>(fine, coarse, cl, fx, fy, fz, x, y)
                                  : samples_t<float, float>(fine, coarse, cl, fx, fy, fz, x, y);
  if (rc) return rc;
  PLL_HIP_TRY(hipStreamSynchronize(coarse->stream));
  return PLL_OK;
}

int pll_closure_create(uint64_t seed, pll_closure** out) {
  if (!out) return fail(PLL_ERR_ARG, "null output pointer");
  *out = nullptr;
  pll_closure* c = new pll_closure();
  auto bail = [&](int rc) {
    pll_closure_destroy(c);
    return rc;
  };
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) return bail(hip_fail(e, "hipGetDevice"));
  e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) return bail(hip_fail(e, "closure stream"));
  e = hipEventCreateWithFlags(&c->ready, hipEventDisableTiming);
  if (e != hipSuccess) return bail(hip_fail(e, "closure event"));
  const size_t pb = (size_t)kNParam * sizeof(float);
  if (hipMalloc(&c->p, pb) != hipSuccess || hipMalloc(&c->m1, pb) != hipSuccess ||
      hipMalloc(&c->m2, pb) != hipSuccess || hipMalloc(&c->grad, pb + sizeof(float)) != hipSuccess ||
      hipMalloc(&c->chunks, (size_t)(kMaxWaves / kRedChunk) * (pb + sizeof(float))) != hipSuccess ||
      hipMalloc(&c->wb, (size_t)kWbLen * sizeof(uint16_t)) != hipSuccess)
    return bail(fail(PLL_ERR_NOMEM, "closure parameters"));
  // Glorot-uniform weights, zero biases (the output bias starts at softplus^-1(1))
  std::vector<float> h(kNParam, 0.f);
  std::mt19937_64 rng(seed);
  auto init = [&](int off, int fan_in, int fan_out, int rows, int cols) {
    const float a = std::sqrt(6.f / (float)(fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (int r = 0; r < rows; ++r)
      for (int q = 0; q < cols; ++q) h[off + r * cols + q] = r < fan_in ? u(rng) : 0.f;
  };
  init(oW1, kIn, kH, kInP, kH);
  init(oW2, kH, kH, kH, kH);
  init(oW3, kH, 1, kH, 1);
  h[oB3] = 0.5413249f;  // log(e - 1)
  std::vector<float> z(kNParam, 0.f);
  if (hipMemcpy(c->p, h.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->m1, z.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->m2, z.data(), pb, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(PLL_ERR_HIP, "closure initialisation"));
  hipLaunchKernelGGL(k_refresh_bf16, dim3((kNParam + 255) / 256), dim3(256), 0, c->stream, c->p,
                     c->wb);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(PLL_ERR_HIP, "closure init"));
  *out = c;
  return PLL_OK;
}

int pll_closure_destroy(pll_closure* c) {
  if (!c) return PLL_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* bufs[] = {c->p, c->m1, c->m2, c->wb, c->partial, c->grad, c->chunks};
  for (void* q : bufs)
    if (q) (void)hipFree(q);
  if (c->ready) (void)hipEventDestroy(c->ready);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PLL_OK;
}

int pll_closure_num_params(void) { return kNParam; }

int pll_closure_get_params(pll_closure* c, float* host) {
  if (!c || !host) return fail(PLL_ERR_ARG, "null argument");
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  PLL_HIP_TRY(hipMemcpy(host, c->p, (size_t)kNParam * sizeof(float), hipMemcpyDeviceToHost));
  return PLL_OK;
}

int pll_closure_set_params(pll_closure* c, const float* host) {
  if (!c || !host) return fail(PLL_ERR_ARG, "null argument");
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  PLL_HIP_TRY(hipMemcpy(c->p, host, (size_t)kNParam * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_refresh_bf16, dim3((kNParam + 255) / 256), dim3(256), 0, c->stream, c->p,
                     c->wb);
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

int pll_closure_gradient_noalloc(pll_closure* c, const float* x, const float* y, long n,
                                 float* grad_host, double* loss) {
  if (!c || !x || !y || n <= 0) return fail(PLL_ERR_ARG, "bad arguments");
  int dev = -1;
  PLL_HIP_TRY(hipGetDevice(&dev));
  if (dev != c->device)
    return fail(PLL_ERR_STATE, "pll_closure_gradient_noalloc: the closure's device must be current");
  const long nw = waves_for(n), nblk = nw / kWaves;
  const long nchunk = (nblk + kRedChunk - 1) / kRedChunk;
  if (nblk > kScratchPartials || nchunk > kScratchChunks)
    return fail(PLL_ERR_ARG, "pll_closure_gradient_noalloc: the training grid exceeds the scratch "
                             "array (at most kMaxWaves waves)");
  float* base = nullptr;
  int rc = scratch_on(dev, &base);
  if (rc) return rc;
  float* partial = base;
  float* chunks = base + kScratchPartials * (kNParam + 1);
  float* grad = chunks + kScratchChunks * (kNParam + 1);
  // the kernels of pll_closure_gradient, on the scratch array
  hipLaunchKernelGGL((k_mlp<true>), dim3((unsigned)nblk), dim3(64 * kWaves), 0, c->stream,
                     weights(c), x, y, n, static_cast<float*>(nullptr), partial,
                     tiles_per_wave(n, nw));
  hipLaunchKernelGGL(k_grad_reduce, dim3((kNParam + 1 + 255) / 256, (unsigned)nchunk), dim3(256), 0,
                     c->stream, partial, nblk, chunks);
  hipLaunchKernelGGL(k_grad_final, dim3((kNParam + 1 + 255) / 256), dim3(256), 0, c->stream, chunks,
                     nchunk, grad);
  PLL_HIP_TRY(hipGetLastError());
  float lsum = 0.f;
  if (grad_host)
    PLL_HIP_TRY(hipMemcpyAsync(grad_host, grad, (size_t)kNParam * sizeof(float),
                               hipMemcpyDeviceToHost, c->stream));
  PLL_HIP_TRY(hipMemcpyAsync(&lsum, grad + kNParam, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  if (loss) *loss = (double)lsum / (double)n;
  return PLL_OK;
}

int pll_closure_gradient(pll_closure* c, const float* x, const float* y, long n, float* grad_host,
                         double* loss) {
  if (!c || !x || !y || n <= 0) return fail(PLL_ERR_ARG, "bad arguments");
  int rc = gradient(c, x, y, n);
  if (rc) return rc;
  std::vector<float> g(kNParam + 1);
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  PLL_HIP_TRY(hipMemcpy(g.data(), c->grad, g.size() * sizeof(float), hipMemcpyDeviceToHost));
  if (grad_host) std::copy(g.begin(), g.begin() + kNParam, grad_host);
  if (loss) *loss = (double)g[kNParam] / (double)n;
  return PLL_OK;
}

int pll_closure_train_step(pll_closure* c, const float* x, const float* y, long n, double lr) {
  if (!c || !x || !y || n <= 0 || !(lr > 0)) return fail(PLL_ERR_ARG, "bad arguments");
  int rc = PLL_OK;
  const long nchunk = gradient_chunks(c, x, y, n, false, &rc);
  if (rc) return rc;
  c->step += 1;
  const double b1 = 0.9, b2 = 0.999;
  const float c1 = (float)(1.0 - std::pow(b1, (double)c->step));
  const float c2 = (float)(1.0 - std::pow(b2, (double)c->step));
  hipLaunchKernelGGL(k_adam, dim3((kNParam + 1 + 255) / 256), dim3(256), 0, c->stream, c->p,
                     c->m1, c->m2, c->chunks, nchunk, c->grad, (float)lr, (float)b1, (float)b2,
                     1e-8f, c1, c2, c->wb);
  PLL_HIP_TRY(hipGetLastError());
  c->version += 1;
  return PLL_OK;
}

int pll_closure_predict(pll_closure* c, const float* x, long n, float* out) {
  if (!c || !x || !out || n <= 0) return fail(PLL_ERR_ARG, "bad arguments");
  const long nw = waves_for(n);
  hipLaunchKernelGGL((k_mlp<false>), dim3((unsigned)(nw / kWaves)), dim3(64 * kWaves), 0, c->stream,
                     weights(c), x, static_cast<const float*>(nullptr), n, out,
                     static_cast<float*>(nullptr), tiles_per_wave(n, nw));
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

int pll_closure_synchronize(pll_closure* c) {
  if (!c) return fail(PLL_ERR_ARG, "null closure");
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

}  // extern "C"
