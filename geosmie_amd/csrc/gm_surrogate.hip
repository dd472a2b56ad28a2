// Query handling of the LUT surrogate: feature transform, hull/trust gate and
// output transform.  No counterpart in the reference.
//
// A query is q = (ln r_eff, sigma_g, m_r, ln m_i, lambda, f_w): r_eff [um] the
// wet effective radius of a lognormal number distribution,
// r_eff = gf r_mode exp(2.5 ln^2 sigma_g); f_w = 1 - gf^-3 the water volume
// fraction of the humidity growth (dointegration.py:497-507); m_r + i m_i the
// dry refractive index; lambda [um].
//   network features  (ln x_eff, sigma_g, Re m_mix, ln Im m_mix, ln(gf / lambda), f_w),
//                     standardised: the effective size parameter 2 pi r_eff / lambda and
//                     the volume-mixed index m_mix = m_w + (m - m_w)(1 - f_w) that the
//                     humidity rule gives (dointegration.py:497-507), gf / lambda locating
//                     the radius grid's end in size-parameter space, and f_w;
//   gate coordinates  (ln r_mode, sigma_g, m_r, ln m_i, ln lambda, f_w) with the
//                     dry mode radius r_mode = r_eff / (gf exp(2.5 ln^2 sigma_g)).
// flags[i] = 0: served by the network; 1: outside the training box [lo, hi];
// 2: inside, but its cell of the uniform n_cells grid over the box is not
// trusted (validation error above the bound).
// Table rows and network targets (before standardisation) correspond as
//   t0 = ln Qext - ln B,  t1 = ln SSA,  t2 = ln max(g, 1e-8),  t_{3+l} = alpha1_l / (2l + 1):
// the squared error of t0..t2 is the squared relative error of Qext, SSA and
// g, the quantities the gate judges.  B is the bulk Rayleigh extinction of
// the query's lognormal, saturated at 2 (models trained without it pass no
// queries and use B = 1):
//   Q_R = 8/3 |K|^2 x_eff^4 exp(6 s^2) + 4 Im(K) x_eff,   B = Q_R / (1 + Q_R / 2),
// K = (m^2 - 1) / (m^2 + 2) at the volume-mixed index, s = ln sigma_g.  These
// are the small-particle efficiencies Q_sca = 8/3 |K|^2 x^4, Q_abs = 4 Im(K) x
// averaged over the area-weighted lognormal: <r^6>/<r^2> = r_eff^4 exp(6 s^2),
// <r^3>/<r^2> = r_eff.  Below x_eff ~ 0.3 bulk Qext spans many decades and
// follows Q_R to O(x^2); the network then only learns the PSD/grid residual.
#include <math.h>
#include <stdint.h>

#include "gm_ctx.h"

namespace gm {

constexpr int kSurrIn = 6;
constexpr float kGFloor = 1e-8f;

struct Cells {
  int n[kSurrIn];
};

__global__ void k_surr_prep(int64_t n, const float* __restrict__ q, const float* __restrict__ fmean,
                            const float* __restrict__ fstd, const float* __restrict__ lo,
                            const float* __restrict__ hi, Cells cells,
                            const uint8_t* __restrict__ cell_ok, float mw_re,
                            float* __restrict__ x, int32_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* qi = q + i * kSurrIn;
  const float llam = logf(qi[4]);
  const float dry = 1.0f - qi[5];                 // gf^-3
  const float lgf = -logf(dry) * (1.0f / 3.0f);   // ln gf
  float f[kSurrIn] = {qi[0] - llam + 1.8378770664093453f,   // ln(2 pi r_eff / lambda)
                      qi[1], mw_re + (qi[2] - mw_re) * dry, qi[3] + logf(dry), lgf - llam, qi[5]};
#pragma unroll
  for (int j = 0; j < kSurrIn; ++j) x[i * kSurrIn + j] = (f[j] - fmean[j]) / fstd[j];
  const float ls = logf(qi[1]);
  float gc[kSurrIn] = {qi[0] - lgf - 2.5f * ls * ls, qi[1], qi[2], qi[3], llam, qi[5]};
  int flag = 0;
#pragma unroll
  for (int j = 0; j < kSurrIn; ++j)
    if (!(gc[j] >= lo[j] && gc[j] <= hi[j])) flag = 1;   // NaN lands here too
  if (flag == 0) {
    int64_t cell = 0;
#pragma unroll
    for (int j = 0; j < kSurrIn; ++j) {
      const float span = hi[j] - lo[j];
      int c = span > 0.0f ? (int)((gc[j] - lo[j]) / span * (float)cells.n[j]) : 0;
      c = min(max(c, 0), cells.n[j] - 1);
      cell = cell * cells.n[j] + c;
    }
    if (!cell_ok[cell]) flag = 2;
  }
  flags[i] = flag;
}

// ln B of one query (see the header comment); FP32 throughout, log domain.
__device__ __forceinline__ float rayleigh_log_b(const float* __restrict__ qi, float mw_re) {
  const float dry = 1.0f - qi[5];
  const float lx = qi[0] - logf(qi[4]) + 1.8378770664093453f;   // ln x_eff
  const float ls = logf(qi[1]);
  const float mr = mw_re + (qi[2] - mw_re) * dry, mi = expf(qi[3]) * dry;
  const float a = mr * mr - mi * mi, b = 2.0f * mr * mi;          // m^2 = a + i b
  const float den = (a + 2.0f) * (a + 2.0f) + b * b;
  const float k2 = ((a - 1.0f) * (a - 1.0f) + b * b) / den;      // |K|^2
  const float ln_sca = 0.98082925301172f + logf(k2) + 4.0f * lx + 6.0f * ls * ls;   // ln 8/3
  const float ln_abs = 1.38629436111989f + logf(3.0f * b / den) + lx;               // ln 4
  const float hi = fmaxf(ln_sca, ln_abs), lo = fminf(ln_sca, ln_abs);
  const float lq = hi + log1pf(expf(lo - hi));                   // ln Q_R
  const float u = lq - 0.69314718055995f;                        // ln(Q_R / 2)
  return lq - (fmaxf(u, 0.0f) + log1pf(expf(-fabsf(u))));        // ln Q_R - ln(1 + Q_R/2)
}

__global__ void k_surr_post(int64_t n, int n_out, const float* __restrict__ q, float mw_re,
                            const float* __restrict__ y, const float* __restrict__ tmean,
                            const float* __restrict__ tstd, float* __restrict__ rows) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n_out) return;
  const int j = (int)(e % n_out);
  const float t = y[e] * tstd[j] + tmean[j];
  float v;
  if (j == 0) v = expf(q ? t + rayleigh_log_b(q + (e / n_out) * kSurrIn, mw_re) : t);
  else if (j == 1) v = fminf(expf(t), 1.0f);
  else if (j == 2) v = expf(t);
  else v = (float)(2 * (j - 3) + 1) * t;
  rows[e] = v;
}

__global__ void k_surr_targets(int64_t n, int n_out, const float* __restrict__ q, float mw_re,
                               const float* __restrict__ rows, const float* __restrict__ tmean,
                               const float* __restrict__ tstd, float* __restrict__ y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n_out) return;
  const int j = (int)(e % n_out);
  const float v = rows[e];
  float t;
  if (j == 0) t = q ? logf(v) - rayleigh_log_b(q + (e / n_out) * kSurrIn, mw_re) : logf(v);
  else if (j == 1) t = logf(v);
  else if (j == 2) t = logf(fmaxf(v, kGFloor));
  else t = v / (float)(2 * (j - 3) + 1);
  y[e] = (t - tmean[j]) / tstd[j];
}

}  // namespace gm

extern "C" {

// See include/geosmie.h.
int gm_surrogate_prep(gm_ctx* ctx, int64_t n, const float* q, const float* feat_mean,
                      const float* feat_std, const float* box_lo, const float* box_hi,
                      const int* n_cells_host, const uint8_t* cell_ok, int64_t n_cell_ok,
                      double m_water_re, float* x, int32_t* flags, gm_stream stream) {
  if (!n_cells_host) return gm_fail(GM_EINVAL, "gm_surrogate_prep: n_cells is NULL");
  gm::Cells cells;
  int64_t total = 1;
  for (int j = 0; j < gm::kSurrIn; ++j) {
    if (n_cells_host[j] < 1 || n_cells_host[j] > (1 << 16))
      return gm_fail(GM_EINVAL, "gm_surrogate_prep: n_cells must be in [1, 65536]");
    cells.n[j] = n_cells_host[j];
    total *= n_cells_host[j];
    if (total > ((int64_t)1 << 31)) return gm_fail(GM_EINVAL, "gm_surrogate_prep: too many cells");
  }
  // every cell index the kernel can form must lie inside cell_ok
  if (n_cell_ok != total)
    return gm_fail(GM_EINVAL, "gm_surrogate_prep: cell_ok has %lld entries, the cell grid %lld",
                   (long long)n_cell_ok, (long long)total);
  if (!ctx || n < 0) return gm_fail(GM_EINVAL, "gm_surrogate_prep: bad arguments");
  if (n == 0) return GM_OK;
  if (!q || !feat_mean || !feat_std || !box_lo || !box_hi || !cell_ok || !x || !flags)
    return gm_fail(GM_EINVAL, "gm_surrogate_prep: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(gm::k_surr_prep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, q, feat_mean, feat_std, box_lo, box_hi, cells,
                     cell_ok, (float)m_water_re, x, flags);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "k_surr_prep: %s", hipGetErrorString(e));
  return GM_OK;
}

int gm_surrogate_post(gm_ctx* ctx, int64_t n, int n_out, const float* q, double m_water_re,
                      const float* y, const float* tgt_mean, const float* tgt_std, float* rows,
                      gm_stream stream) {
  if (!ctx || n < 0 || n_out < 3) return gm_fail(GM_EINVAL, "gm_surrogate_post: bad arguments");
  if (n == 0) return GM_OK;
  if (!y || !tgt_mean || !tgt_std || !rows) return gm_fail(GM_EINVAL, "gm_surrogate_post: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  const int64_t tot = n * n_out;
  hipLaunchKernelGGL(gm::k_surr_post, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, n_out, q, (float)m_water_re, y, tgt_mean, tgt_std,
                     rows);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "k_surr_post: %s", hipGetErrorString(e));
  return GM_OK;
}

int gm_surrogate_targets(gm_ctx* ctx, int64_t n, int n_out, const float* q, double m_water_re,
                         const float* rows, const float* tgt_mean, const float* tgt_std, float* y,
                         gm_stream stream) {
  if (!ctx || n < 0 || n_out < 3) return gm_fail(GM_EINVAL, "gm_surrogate_targets: bad arguments");
  if (n == 0) return GM_OK;
  if (!rows || !tgt_mean || !tgt_std || !y)
    return gm_fail(GM_EINVAL, "gm_surrogate_targets: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  const int64_t tot = n * n_out;
  hipLaunchKernelGGL(gm::k_surr_targets, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, n_out, q, (float)m_water_re, rows, tgt_mean,
                     tgt_std, y);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "k_surr_targets: %s", hipGetErrorString(e));
  return GM_OK;
}

}  // extern "C"
