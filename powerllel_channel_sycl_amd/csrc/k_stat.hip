// k_stat.hip -- run-time statistics on the device: the per-point sums that
// calcStat keeps (src/mod_statistics.cc:30-106), sampled with cell-centred
// velocities
//   uc = (u(i-1)+u(i))/2 + u_crf,  vc = (v(j-1)+v(j))/2,  wc = (w(k-1)+w(k))/2
// and the quantities, in the reference's order (src/mod_dataIO.cc:39-63):
//   u v w u2 v2 w2 uv uw vw p p2.
// Sums are FP64 in both precisions (FP32 fields are widened on load), so the
// mean profile and Reynolds stresses of a long run do not drift with the
// sample count.  Plane sums of a selected quantity reduce in a fixed order.
#include "pll_internal.hpp"

namespace pll {
namespace {

constexpr int kStTx = 64, kStTy = 4;

struct StatSums {
  double* s[11];
};

template <typename T>
__global__ __launch_bounds__(kStTx* kStTy) void k_stat_accum(Geom g, const T* __restrict__ u,
                                                            const T* __restrict__ v,
                                                            const T* __restrict__ w,
                                                            const T* __restrict__ p,
                                                            double u_crf, StatSums st, int kchunk) {
  const int i = blockIdx.x * kStTx + threadIdx.x + 1;
  const int j = blockIdx.y * kStTy + threadIdx.y + 1;
  if (i > g.nx || j > g.ny) return;
  const int k0 = blockIdx.z * kchunk + 1;
  const int k1 = min(k0 + kchunk - 1, g.nz);
  const long nxy = (long)g.nx * g.ny;
  double wlo = (double)w[g.idx(i, j, k0 - 1)];
  for (int k = k0; k <= k1; ++k) {
    const long c = g.idx(i, j, k);
    const double whi = (double)w[c];
    const double uc = ((double)u[c - 1] + (double)u[c]) * 0.5 + u_crf;
    const double vc = ((double)v[c - g.sx] + (double)v[c]) * 0.5;
    const double wc = (wlo + whi) * 0.5;
    wlo = whi;
    const double pc = (double)p[c];
    const long o = (i - 1) + (long)g.nx * (j - 1) + nxy * (k - 1);
    const double q[11] = {uc, vc, wc, uc * uc, vc * vc, wc * wc, uc * vc, uc * wc, vc * wc, pc, pc * pc};
#pragma unroll
    for (int t = 0; t < 11; ++t)
      if (st.s[t]) st.s[t][o] += q[t];
  }
}

// out[k] = sum over the plane k of sum[] (one block per plane, fixed order).
__global__ __launch_bounds__(256) void k_plane_sum(const double* __restrict__ sum, int nxy,
                                                   double* __restrict__ out) {
  __shared__ double red[256];
  const double* plane = sum + (long)blockIdx.x * nxy;
  double acc = 0.0;
  for (int t = threadIdx.x; t < nxy; t += 256) acc += plane[t];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <typename T>
int accum_t(pll_context* c) {
  StatSums st;
  for (int t = 0; t < 11; ++t) st.s[t] = c->d_stat[t];
  const int kc = c->g.nz < 16 ? c->g.nz : 16;
  const dim3 grid((c->g.nx + kStTx - 1) / kStTx, (c->g.ny + kStTy - 1) / kStTy,
                  (c->g.nz + kc - 1) / kc);
  TimerScope ts(c, K_MISC);
  hipLaunchKernelGGL((k_stat_accum<T>), grid, dim3(kStTx, kStTy), 0, c->stream, c->g,
                     static_cast<const T*>(c->fld[PLL_U]), static_cast<const T*>(c->fld[PLL_V]),
                     static_cast<const T*>(c->fld[PLL_W]), static_cast<const T*>(c->fld[PLL_P]),
                     c->u_crf, st, kc);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

int check_var(pll_context* c, int var) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  if (var < 0 || var > 10) return fail(PLL_ERR_ARG, "statistic index must be in 0..10");
  if (!c->d_stat[var]) return fail(PLL_ERR_STATE, "statistic not selected by pll_stat_enable");
  return PLL_OK;
}

}  // namespace
}  // namespace pll

using namespace pll;

extern "C" {

int pll_stat_enable(pll_context* c, const int which[11]) {
  if (!c || !which) return fail(PLL_ERR_ARG, "null argument");
  const size_t bytes = (size_t)c->nspec * sizeof(double);
  for (int t = 0; t < 11; ++t) {
    if (which[t] && !c->d_stat[t]) {
      hipError_t e = ctx_malloc(c, &c->d_stat[t], bytes);
      if (e != hipSuccess) {
        c->d_stat[t] = nullptr;
        return fail(PLL_ERR_NOMEM, "statistics sums do not fit in device memory");
      }
    } else if (!which[t] && c->d_stat[t]) {
      PLL_HIP_TRY(hipStreamSynchronize(c->stream));
      (void)hipFree(c->d_stat[t]);
      c->d_stat[t] = nullptr;
      c->dev_bytes -= bytes;
    }
    if (c->d_stat[t]) PLL_HIP_TRY(hipMemsetAsync(c->d_stat[t], 0, bytes, c->stream));
  }
  return PLL_OK;
}

int pll_stat_accumulate(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  bool any = false;
  for (double* p : c->d_stat) any |= p != nullptr;
  if (!any) return fail(PLL_ERR_STATE, "no statistic selected by pll_stat_enable");
  return c->prec == PLL_FP64 ? accum_t<double>(c) : accum_t<float>(c);
}

int pll_stat_download(pll_context* c, int var, double* host) {
  int rc = check_var(c, var);
  if (rc) return rc;
  if (!host) return fail(PLL_ERR_ARG, "null host pointer");
  PLL_HIP_TRY(hipMemcpyAsync(host, c->d_stat[var], (size_t)c->nspec * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

int pll_stat_upload(pll_context* c, int var, const double* host) {
  int rc = check_var(c, var);
  if (rc) return rc;
  if (!host) return fail(PLL_ERR_ARG, "null host pointer");
  PLL_HIP_TRY(hipMemcpyAsync(c->d_stat[var], host, (size_t)c->nspec * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

int pll_stat_plane_sums(pll_context* c, int var, double* host) {
  int rc = check_var(c, var);
  if (rc) return rc;
  if (!host) return fail(PLL_ERR_ARG, "null host pointer");
  if (c->npartial < c->g.nz) return fail(PLL_ERR_STATE, "reduction scratch too small");
  hipLaunchKernelGGL(k_plane_sum, dim3(c->g.nz), dim3(256), 0, c->stream, c->d_stat[var],
                     c->g.nx * c->g.ny, c->d_partial);
  PLL_HIP_TRY(hipGetLastError());
  PLL_HIP_TRY(hipMemcpyAsync(host, c->d_partial, (size_t)c->g.nz * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

}  // extern "C"
