// k_closure.hip -- the learned eddy-viscosity closure: a 10 -> 64 -> 64 -> 1
// multilayer perceptron (tanh hidden layers, softplus output) with BF16
// weights and FP32 accumulation on the MFMA units of gfx950, its training
// (forward, backward, Adam) and its inference.  The reference has no closure.
//
// Tiles.  A wave owns 32 samples at a time.  The hidden layers are
// v_mfma_f32_32x32x16_bf16 products: A is 32 x 16 (lane l holds row l%32,
// k = 8*(l/32) .. +7), B is 16 x 32 (lane l holds column l%32, the same k),
// D is 32 x 32 FP32 (lane l holds column l%32, rows 8*(j/4) + 4*(l/32) + j%4
// for j = 0..15).  Activations pass between layers through LDS as BF16 (the
// A operand of the next product); the output layer (one column) is a
// per-sample FP32 dot product.  Rounding to BF16 is round-to-nearest-even at
// fixed points -- the inputs, the weights, the activations h1 and h2 and the
// back-propagated deltas -- so a CPU restatement can follow it exactly.
//
// Training is deterministic: each wave accumulates its gradient over its
// tiles in registers, writes one partial per wave, and a second kernel sums
// the partials in a fixed order before the Adam update.
#include <cstdint>
#include <random>
#include <vector>

#include "pll_internal.hpp"

namespace pll {
namespace {

constexpr int kIn = 10;   // features
constexpr int kInP = 16;  // padded to one MFMA K block
constexpr int kH = 64;    // hidden width
// parameter vector (FP32 master): W1[16][64], b1[64], W2[64][64], b2[64], W3[64], b3
constexpr int oW1 = 0, oB1 = oW1 + kInP * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH, kNParam = oB3 + 1;
constexpr int kWaves = 4;  // waves per block
constexpr int kTilesPerWave = 4;

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__host__ __device__ inline uint16_t to_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                  // nearest, ties to even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  memcpy(&x, &u, 4);
  return x;
}

__device__ inline float softplus(float o) { return o > 20.f ? o : log1pf(expf(o)); }
__device__ inline float sigmoid(float o) { return 1.f / (1.f + expf(-o)); }

struct Mlp {
  const float* p;      // FP32 parameters (biases and output weights are read from here)
  const uint16_t* w1;  // BF16 W1 [16][64]
  const uint16_t* w2;  // BF16 W2 [64][64]
  const uint16_t* w3;  // BF16 W3 [64]
};

// Per-wave LDS: X (32 x 16), h1 and h2 (32 x 64), deltas (32 x 64), all BF16.
struct WaveLds {
  uint16_t x[32][kInP];
  uint16_t h1[32][kH];
  uint16_t h2[32][kH];
  uint16_t d2[32][kH];
  uint16_t d1[32][kH];
  float g[32];    // dL/do per sample
  float red[64];  // lane exchange
};

__device__ inline bf16x8 load8(const uint16_t* p) {
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (short)p[e];
  return v;
}

// D-layout row of accumulator element j for this lane.
__device__ inline int drow(int j, int lane) { return 8 * (j / 4) + 4 * (lane / 32) + (j % 4); }

// Forward pass of one 32-sample tile.  x: features of samples m0..m0+31 (rows
// beyond n are zero); fills L.x/h1/h2 and returns z1/z2 activations in D layout.
__device__ inline void forward_tile(const Mlp& w, const float* __restrict__ x, long m0, long n,
                                    WaveLds& L, int lane, f32x16 h1[2], f32x16 h2[2],
                                    float& out_o) {
  // stage X (BF16, 10 features + 6 zeros)
  for (int e = lane; e < 32 * kInP; e += 64) {
    const int m = e / kInP, k = e % kInP;
    const long s = m0 + m;
    L.x[m][k] = to_bf16((k < kIn && s < n) ? x[s * kIn + k] : 0.f);
  }
  __syncthreads();
  const int col = lane % 32, kh = 8 * (lane / 32);
  // layer 1: h1 = tanh(X W1 + b1)
  const bf16x8 a1 = load8(&L.x[col][kh]);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    bf16x8 b1;
#pragma unroll
    for (int e = 0; e < 8; ++e) b1[e] = (short)w.w1[(kh + e) * kH + 32 * t + col];
    f32x16 acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
    const float bias = w.p[oB1 + 32 * t + col];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float h = tanhf(acc[j] + bias);
      acc[j] = h;
      L.h1[drow(j, lane)][32 * t + col] = to_bf16(h);
    }
    h1[t] = acc;
  }
  __syncthreads();
  // layer 2: h2 = tanh(h1 W2 + b2)
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    f32x16 acc = {};
#pragma unroll
    for (int kb = 0; kb < kH / 16; ++kb) {
      const bf16x8 a = load8(&L.h1[col][16 * kb + kh]);
      bf16x8 b;
#pragma unroll
      for (int e = 0; e < 8; ++e) b[e] = (short)w.w2[(16 * kb + kh + e) * kH + 32 * t + col];
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const float bias = w.p[oB2 + 32 * t + col];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float h = tanhf(acc[j] + bias);
      acc[j] = h;
      L.h2[drow(j, lane)][32 * t + col] = to_bf16(h);
    }
    h2[t] = acc;
  }
  __syncthreads();
  // output: o = h2 . W3 + b3 (lanes 0..31, one sample each, fixed order)
  out_o = 0.f;
  if (lane < 32) {
    float o = 0.f;
    for (int k = 0; k < kH; ++k) o += from_bf16(L.h2[lane][k]) * from_bf16(w.w3[k]);
    out_o = o + w.p[oB3];
  }
}

// Training: one partial gradient (and loss sum) per wave.  Loss = mean over
// the n samples of (softplus(o) - y)^2.
template <bool TRAIN>
__global__ __launch_bounds__(64 * kWaves) void k_mlp(Mlp w, const float* __restrict__ x,
                                                      const float* __restrict__ y, long n,
                                                      float* __restrict__ out,
                                                      float* __restrict__ partial) {
  __shared__ WaveLds lds[kWaves];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  WaveLds& L = lds[wave];
  const long gw = (long)blockIdx.x * kWaves + wave;  // global wave index
  const int col = lane % 32, kh = 8 * (lane / 32);
  f32x16 gW2[2][2] = {};  // dW2 tiles [k-tile][n-tile] (rows k, cols n)
  f32x16 gW1[2] = {};     // dW1 tiles [n-tile] (rows i < 16 meaningful)
  float gb1[2] = {0.f, 0.f}, gb2[2] = {0.f, 0.f}, gw3[2] = {0.f, 0.f}, gb3 = 0.f, loss = 0.f;
  const float inv_n = 1.f / (float)n;
  for (int tile = 0; tile < kTilesPerWave; ++tile) {
    const long m0 = (gw * kTilesPerWave + tile) * 32;
    // every wave of a block runs the same number of tiles (barriers inside)
    f32x16 h1[2], h2[2];
    float o;
    forward_tile(w, x, m0, n, L, lane, h1, h2, o);
    if (!TRAIN) {
      if (lane < 32 && m0 + lane < n) out[m0 + lane] = softplus(o);
      __syncthreads();
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
    __syncthreads();
    // output layer gradients and delta2 = g w3 (1 - h2^2)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int nn = 32 * t + col;
      const float w3n = from_bf16(w.w3[nn]);
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int m = drow(j, lane);
        const float g = L.g[m];
        gw3[t] += from_bf16(L.h2[m][nn]) * g;
        const float h = h2[t][j];
        const float d2 = g * w3n * (1.f - h * h);
        gb2[t] += d2;
        L.d2[m][nn] = to_bf16(d2);
      }
    }
    if (lane == 0)
      for (int m = 0; m < 32; ++m) gb3 += L.g[m];
    __syncthreads();
    // dW2 += h1^T d2 : rows k (h1 column), cols n, K = samples
#pragma unroll
    for (int tk = 0; tk < 2; ++tk)
#pragma unroll
      for (int tn = 0; tn < 2; ++tn)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          bf16x8 a, b;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            a[e] = (short)L.h1[16 * ks + kh + e][32 * tk + col];
            b[e] = (short)L.d2[16 * ks + kh + e][32 * tn + col];
          }
          gW2[tk][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, gW2[tk][tn], 0, 0, 0);
        }
    // delta1 = (d2 W2^T) (1 - h1^2) : rows m, cols k, K = n
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
        const float d1 = acc[j] * (1.f - h * h);
        gb1[t] += d1;
        L.d1[drow(j, lane)][32 * t + col] = to_bf16(d1);
      }
    }
    __syncthreads();
    // dW1 += X^T d1 : rows i (< 16), cols k, K = samples
#pragma unroll
    for (int tn = 0; tn < 2; ++tn)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          a[e] = col < kInP ? (short)L.x[16 * ks + kh + e][col] : (short)0;
          b[e] = (short)L.d1[16 * ks + kh + e][32 * tn + col];
        }
        gW1[tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, gW1[tn], 0, 0, 0);
      }
    __syncthreads();
  }
  if (!TRAIN) return;
  // write this wave's partial: lanes l and l+32 hold different rows of the same
  // columns; the per-column sums (biases, W3) are combined through LDS
  float* P = partial + gw * (kNParam + 1);
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
    __syncthreads();
    if (lane < 32) P[off + 32 * t + lane] = L.red[lane] + L.red[lane + 32];
    __syncthreads();
  };
  for (int t = 0; t < 2; ++t) {
    colsum(gb1[t], t, oB1);
    colsum(gb2[t], t, oB2);
    colsum(gw3[t], t, oW3);
  }
  L.red[lane] = lane < 32 ? loss : 0.f;
  __syncthreads();
  if (lane == 0) {
    float l = 0.f;
    for (int q = 0; q < 32; ++q) l += L.red[q];
    P[oB3] = gb3;
    P[kNParam] = l;
  }
}

// Sum of the per-wave partials in wave order: grad[i] for i < kNParam, loss sum last.
__global__ __launch_bounds__(256) void k_grad_reduce(const float* __restrict__ partial,
                                                     long nwaves, float* __restrict__ grad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i > kNParam) return;
  float s = 0.f;
  for (long q = 0; q < nwaves; ++q) s += partial[q * (kNParam + 1) + i];
  grad[i] = s;
}

// Adam (Kingma & Ba, 2015) on the FP32 master parameters; refreshes the BF16
// weights.  The padded rows 10..15 of W1 stay zero.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, float* __restrict__ m1,
                                              float* __restrict__ m2, const float* __restrict__ g,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2, uint16_t* __restrict__ w1b,
                                              uint16_t* __restrict__ w2b,
                                              uint16_t* __restrict__ w3b) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kNParam) return;
  const bool pad = i < oB1 && (i / kH) >= kIn;
  float v = p[i];
  if (!pad) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    v -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
  }
  if (i < oB1) w1b[i - oW1] = to_bf16(pad ? 0.f : v);
  else if (i >= oW2 && i < oB2) w2b[i - oW2] = to_bf16(v);
  else if (i >= oW3 && i < oB3) w3b[i - oW3] = to_bf16(v);
}

__global__ void k_refresh_bf16(const float* __restrict__ p, uint16_t* __restrict__ w1b,
                               uint16_t* __restrict__ w2b, uint16_t* __restrict__ w3b) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kNParam) return;
  if (i < oB1) w1b[i - oW1] = to_bf16(((i / kH) >= kIn) ? 0.f : p[i]);
  else if (i >= oW2 && i < oB2) w2b[i - oW2] = to_bf16(p[i]);
  else if (i >= oW3 && i < oB3) w3b[i - oW3] = to_bf16(p[i]);
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
  long npartial_waves = 0;
  float* grad = nullptr;  // kNParam + 1 (loss sum last)
  long step = 0;
  hipStream_t stream = nullptr;
};

namespace {

Mlp weights(const pll_closure* c) {
  return Mlp{c->p, c->wb, c->wb + kInP * kH, c->wb + kInP * kH + kH * kH};
}

int ensure_partial(pll_closure* c, long nwaves) {
  if (nwaves <= c->npartial_waves) return PLL_OK;
  if (c->partial) (void)hipFree(c->partial);
  c->partial = nullptr;
  PLL_HIP_TRY(hipMalloc(&c->partial, (size_t)nwaves * (kNParam + 1) * sizeof(float)));
  c->npartial_waves = nwaves;
  return PLL_OK;
}

long waves_for(long n) {
  const long tiles = (n + 31) / 32;
  const long waves = (tiles + kTilesPerWave - 1) / kTilesPerWave;
  return (waves + kWaves - 1) / kWaves * kWaves;
}

int gradient(pll_closure* c, const float* x, const float* y, long n) {
  const long nw = waves_for(n);
  int rc = ensure_partial(c, nw);
  if (rc) return rc;
  hipLaunchKernelGGL((k_mlp<true>), dim3((unsigned)(nw / kWaves)), dim3(64 * kWaves), 0, c->stream,
                     weights(c), x, y, n, static_cast<float*>(nullptr), c->partial);
  hipLaunchKernelGGL(k_grad_reduce, dim3((kNParam + 1 + 255) / 256), dim3(256), 0, c->stream,
                     c->partial, nw, c->grad);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

}  // namespace

extern "C" {

int pll_closure_create(uint64_t seed, pll_closure** out) {
  if (!out) return fail(PLL_ERR_ARG, "null output pointer");
  *out = nullptr;
  pll_closure* c = new pll_closure();
  auto bail = [&](int rc) {
    pll_closure_destroy(c);
    return rc;
  };
  hipError_t e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) return bail(hip_fail(e, "closure stream"));
  const size_t pb = (size_t)kNParam * sizeof(float);
  if (hipMalloc(&c->p, pb) != hipSuccess || hipMalloc(&c->m1, pb) != hipSuccess ||
      hipMalloc(&c->m2, pb) != hipSuccess || hipMalloc(&c->grad, pb + sizeof(float)) != hipSuccess ||
      hipMalloc(&c->wb, (size_t)(kInP * kH + kH * kH + kH) * sizeof(uint16_t)) != hipSuccess)
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
                     c->wb, c->wb + kInP * kH, c->wb + kInP * kH + kH * kH);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(PLL_ERR_HIP, "closure init"));
  *out = c;
  return PLL_OK;
}

int pll_closure_destroy(pll_closure* c) {
  if (!c) return PLL_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* bufs[] = {c->p, c->m1, c->m2, c->wb, c->partial, c->grad};
  for (void* q : bufs)
    if (q) (void)hipFree(q);
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
                     c->wb, c->wb + kInP * kH, c->wb + kInP * kH + kH * kH);
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
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
  int rc = gradient(c, x, y, n);
  if (rc) return rc;
  c->step += 1;
  const double b1 = 0.9, b2 = 0.999;
  const float c1 = (float)(1.0 - std::pow(b1, (double)c->step));
  const float c2 = (float)(1.0 - std::pow(b2, (double)c->step));
  hipLaunchKernelGGL(k_adam, dim3((kNParam + 255) / 256), dim3(256), 0, c->stream, c->p, c->m1,
                     c->m2, c->grad, (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2, c->wb,
                     c->wb + kInP * kH, c->wb + kInP * kH + kH * kH);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

int pll_closure_predict(pll_closure* c, const float* x, long n, float* out) {
  if (!c || !x || !out || n <= 0) return fail(PLL_ERR_ARG, "bad arguments");
  const long nw = waves_for(n);
  hipLaunchKernelGGL((k_mlp<false>), dim3((unsigned)(nw / kWaves)), dim3(64 * kWaves), 0, c->stream,
                     weights(c), x, static_cast<const float*>(nullptr), n, out,
                     static_cast<float*>(nullptr));
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

int pll_closure_synchronize(pll_closure* c) {
  if (!c) return fail(PLL_ERR_ARG, "null closure");
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

}  // extern "C"
