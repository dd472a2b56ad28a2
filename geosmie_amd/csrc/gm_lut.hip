// LUT sweep on gfx950: bulk optics of lognormal size distributions for many
// (refractive index, wavelength, humidity) slabs at once.
//
// A slab is one (m_dry, lambda, gf) combination over the dry radius grid
// r_dry[n_r] -- one iteration of runoptics' (lambda, rh) loop
// (src/geosmie/dointegration.py:787-889).  Points are slab-major (point
// s * n_r + p).  Humidity growth follows the reference's 'simple' rule
// (dointegration.py:497-507): wet radius gf * r_dry and volume mixing with
// water  m = m_w + (m_dry - m_w) * (1/gf)^3.
//
// For every (slab s, PSD k) entry, with number weights w[k][p] on the dry
// grid (growth moves every particle of a bin together, so the weights carry
// over), integratePSD's area weighting ('dumix', dointegration.py:1102-1190):
//   Qext = sum w qext r^2 / sum w r^2,  Qsca likewise,  SSA = Qsca / Qext (:961),
//   g    = sum w g qsca r^2 / sum w qsca r^2   (the Qsca weight, :1111-1112),
// and the first n_mom Legendre moments of P11 as spher_expan.f forms alpha1
// (SPHER_EXPAN :313-347 with LINTERPOL and the Gauss nodes of GAUSS),
// normalised by alpha1_0 (CNORM, :95-103).  P11's own normalisation
// (dointegration.py:978-985) cancels in that ratio and is skipped.
// Output rows are float32: Qext, SSA, g, alpha1_0 .. alpha1_{n_mom-1}.
#include <math.h>

#include "gm_complex.h"
#include "gm_ctx.h"

namespace gm {

constexpr int kLutFeat = 4;       // r^2, qext r^2, qsca r^2, g qsca r^2
constexpr int kLutMaxMom = 64;

__global__ void k_lut_points(int64_t n_slab, int n_r, const double* __restrict__ r_dry,
                             const double* __restrict__ lam, const double* __restrict__ m_dry,
                             const double* __restrict__ gf, double mw_re, double mw_im,
                             double* __restrict__ x, double* __restrict__ eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_slab * n_r) return;
  const int64_t s = i / n_r;
  const int p = (int)(i - s * n_r);
  const double g = gf[s];
  const double rrat = 1.0 / g;
  const double r3 = rrat * rrat * rrat;
  const double mr = mw_re + (m_dry[2 * s] - mw_re) * r3;
  const double mi = mw_im + (m_dry[2 * s + 1] - mw_im) * r3;
  x[i] = 2.0 * 3.14159265358979323846 * g * r_dry[p] / lam[s];
  eps[2 * i] = mr * mr - mi * mi;
  eps[2 * i + 1] = 2.0 * mr * mi;
}

// sums[s][k][c], c < n_ang: sum_p w[k][p] P11[s][p][c];  c = n_ang + f: the
// four area-weighted features.  Per slab this is the product
//   [PSD rows x n_r radii] x [n_r x (n_ang + 4) columns]
// on the FP64 matrix cores (v_mfma_f64_16x16x4_f64: A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k, D[g + 4r][j] in lane j + 16g element r).
// One wavefront per (slab, 64 columns, 48 PSD rows): three 16-row tiles x four
// 16-column tiles, walking the radii four at a time; every plane element is
// read once, the weights (shared by the slabs) come from the cache, and each
// sum is formed in one fixed order.
typedef double double4_ __attribute__((ext_vector_type(4)));
constexpr int kLutRowTiles = 3;   // 48 PSD rows per wavefront
__global__ __launch_bounds__(64) void k_lut_contract(
    int64_t n_slab, int n_r, int n_ang, int n_psd, const double* __restrict__ w,
    int64_t w_slab_stride, const double* __restrict__ x, const double* __restrict__ lam,
    const double* __restrict__ props, const double* __restrict__ p11, double* __restrict__ sums) {
  const int ncol = n_ang + kLutFeat;
  const int ncb = (ncol + 63) / 64;
  const int64_t s = blockIdx.x / ncb;
  const int c0 = (int)(blockIdx.x - s * ncb) * 64;
  const int k0 = blockIdx.y * 16 * kLutRowTiles;
  const int lane = threadIdx.x;
  const int i = lane & 15, kq = lane >> 4;
  const int64_t n_all = n_slab * n_r;
  const double* wb = w + s * w_slab_stride;
  const double rscale = lam[s] / (2.0 * 3.14159265358979323846);
  double4_ acc[kLutRowTiles][4];
#pragma unroll
  for (int mt = 0; mt < kLutRowTiles; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[mt][t] = (double4_){0.0, 0.0, 0.0, 0.0};
  // operands of radius p = p0 + kq: the weight of PSD k0 + 16 mt + i and the
  // column c0 + 16 t + i (a P11 value or, past n_ang, a feature of point p)
  auto load = [&](int p0, double* a, double* bv) {
    const int p = p0 + kq;
    const bool pok = p < n_r;
#pragma unroll
    for (int mt = 0; mt < kLutRowTiles; ++mt) {
      const int k = k0 + 16 * mt + i;
      a[mt] = (pok && k < n_psd) ? wb[(size_t)k * n_r + p] : 0.0;
    }
    const int64_t pt = s * n_r + p;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = c0 + 16 * t + i;
      double v = 0.0;
      if (pok && c < n_ang) {
        v = p11[pt * n_ang + c];
      } else if (pok && c < ncol) {
        const double r = x[pt] * rscale, r2 = r * r;
        const int f = c - n_ang;
        const double qext = props[pt], qsca = props[n_all + pt], g = props[4 * n_all + pt];
        v = f == 0 ? r2 : (f == 1 ? qext * r2 : (f == 2 ? qsca * r2 : g * qsca * r2));
      }
      bv[t] = v;
    }
  };
  double a_n[kLutRowTiles], b_n[4];
  load(0, a_n, b_n);
  for (int p0 = 0; p0 < n_r; p0 += 4) {
    double a_c[kLutRowTiles], b_c[4];
#pragma unroll
    for (int mt = 0; mt < kLutRowTiles; ++mt) a_c[mt] = a_n[mt];
#pragma unroll
    for (int t = 0; t < 4; ++t) b_c[t] = b_n[t];
    if (p0 + 4 < n_r) load(p0 + 4, a_n, b_n);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int mt = 0; mt < kLutRowTiles; ++mt)
        acc[mt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_c[mt], b_c[t], acc[mt][t], 0, 0, 0);
  }
  // lane (i, kq) element r of acc[mt][t]: PSD k0 + 16 mt + kq + 4 r, column c0 + 16 t + i
#pragma unroll
  for (int mt = 0; mt < kLutRowTiles; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = c0 + 16 * t + i;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = k0 + 16 * mt + kq + 4 * r;
        if (k < n_psd && c < ncol) sums[((size_t)s * n_psd + k) * ncol + c] = acc[mt][t][r];
      }
    }
}

// One wavefront per entry: lane l forms alpha1_l from the GSF table G1[a][l].
__global__ __launch_bounds__(64) void k_lut_finish(int n_ang, int ng, int n_mom,
                                                   const double* __restrict__ G1,
                                                   const double* __restrict__ sums,
                                                   float* __restrict__ out) {
  const int64_t e = blockIdx.x;
  const int l = threadIdx.x;
  const int ncol = n_ang + kLutFeat;
  const double* S = sums + e * ncol;
  double acc = 0.0;
  if (l < n_mom)
    for (int a = 0; a < n_ang; ++a) acc = fma(G1[(size_t)a * ng + l], S[a], acc);
  const double alpha = acc * ((double)l + 0.5);
  const double a0 = __shfl(alpha, 0);
  float* o = out + e * (3 + n_mom);
  if (l < n_mom) o[3 + l] = (float)(alpha / a0);
  if (l == 0) {
    const double sr2 = S[n_ang], sext = S[n_ang + 1], ssca = S[n_ang + 2], sg = S[n_ang + 3];
    o[0] = (float)(sext / sr2);
    o[1] = (float)(ssca / sext);
    o[2] = (float)(sg / ssca);
  }
}

// w[k][p] = n(r_p; r_mode_k, sigma_k) dr_p normalised to sum 1 over p: the
// lognormal number density of particleparams.py:113-127 times the bin widths,
// as calculatePSD builds a mode (dointegration.py:569-577).  One block per row.
__global__ __launch_bounds__(256) void k_lut_lognormal(int n_r, const double* __restrict__ r,
                                                       const double* __restrict__ dr,
                                                       const double* __restrict__ rmode,
                                                       const double* __restrict__ sigma,
                                                       double* __restrict__ w) {
  __shared__ double red[256];
  const int64_t k = blockIdx.x;
  const double ls = log(sigma[k]), rm = rmode[k];
  const double c = 1.0 / (sqrt(2.0 * 3.14159265358979323846) * ls);
  double* wk = w + k * n_r;
  double s = 0.0;
  for (int p = threadIdx.x; p < n_r; p += 256) {
    const double t = log(r[p] / rm);
    const double v = exp(-t * t / (2.0 * ls * ls)) * c / r[p] * dr[p];
    wk[p] = v;
    s += v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double inv = 1.0 / red[0];
  for (int p = threadIdx.x; p < n_r; p += 256) wk[p] *= inv;
}

// Truncated lognormal rows for runoptics: out[k][col0[k] + p] = n(r_p) dr_p /
// sum, with n zero where r_p >= rmax[k] or r_p <= rmin[k] (getLogNormPSD,
// particleparams.py:113-127, times the widths and normalised as calculatePSD
// does, dointegration.py:569-577).  r may be size parameters (then rmode,
// rmin, rmax are multiplied by 2 pi / lambda).  One block per row; the sum is
// a fixed-order tree.
__global__ __launch_bounds__(256) void k_lognorm_rows(int n_r, const double* __restrict__ r,
                                                      const double* __restrict__ dr,
                                                      const double* __restrict__ par,
                                                      const int64_t* __restrict__ col0,
                                                      double* __restrict__ w, int64_t ldw) {
  __shared__ double red[256];
  const int64_t k = blockIdx.x;
  // par[k] = (rmode, ln sigma, rmin, rmax)
  const double rm = par[4 * k], ls = par[4 * k + 1], lo = par[4 * k + 2], hi = par[4 * k + 3];
  const int64_t c0 = col0[k];
  if (c0 < 0 || c0 > ldw - n_r) return;   // row left untouched (uniform over the block)
  const double c = 1.0 / (sqrt(2.0 * 3.14159265358979323846) * ls);
  double* wk = w + k * ldw + c0;
  double s = 0.0;
  for (int p = threadIdx.x; p < n_r; p += 256) {
    const double rp = r[p];
    double v = 0.0;
    if (rp < hi && rp > lo) {
      const double t = log(rp / rm);
      v = c / rp * exp(-(t * t) / (2.0 * ls * ls)) * dr[p];
    }
    wk[p] = v;
    s += v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double tot = red[0];
  for (int p = threadIdx.x; p < n_r; p += 256) wk[p] = wk[p] / tot;
}

}  // namespace gm

extern "C" {

int gm_lut_lognormal(gm_ctx* ctx, int64_t n_psd, int n_r, const double* r, const double* dr,
                     const double* rmode, const double* sigma, double* w, gm_stream stream) {
  if (!ctx || n_psd < 0 || n_r < 1) return gm_fail(GM_EINVAL, "gm_lut_lognormal: bad arguments");
  if (n_psd == 0) return GM_OK;
  if (!r || !dr || !rmode || !sigma || !w) return gm_fail(GM_EINVAL, "gm_lut_lognormal: NULL array");
  if (n_psd > 0x7fffffffLL) return gm_fail(GM_EINVAL, "gm_lut_lognormal: too many rows");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(gm::k_lut_lognormal, dim3((unsigned)n_psd), dim3(256), 0, (hipStream_t)stream,
                     n_r, r, dr, rmode, sigma, w);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "k_lut_lognormal: %s", hipGetErrorString(e));
  return GM_OK;
}

int gm_lognorm_rows(gm_ctx* ctx, int64_t n_rows, int n_r, const double* r, const double* dr,
                    const double* par, const int64_t* col0, double* w, int64_t ldw,
                    gm_stream stream) {
  if (!ctx || n_rows < 0 || n_r < 1 || ldw < n_r)
    return gm_fail(GM_EINVAL, "gm_lognorm_rows: bad arguments");
  if (n_rows == 0) return GM_OK;
  if (!r || !dr || !par || !col0 || !w) return gm_fail(GM_EINVAL, "gm_lognorm_rows: NULL array");
  if (n_rows > 0x7fffffffLL) return gm_fail(GM_EINVAL, "gm_lognorm_rows: too many rows");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(gm::k_lognorm_rows, dim3((unsigned)n_rows), dim3(256), 0,
                     (hipStream_t)stream, n_r, r, dr, par, col0, w, ldw);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "k_lognorm_rows: %s", hipGetErrorString(e));
  return GM_OK;
}

// See include/geosmie.h.
int gm_lut_points(gm_ctx* ctx, int64_t n_slab, int n_r, const double* r_dry, const double* lam,
                  const double* m_dry, const double* gf, double mw_re, double mw_im, double* x,
                  double* eps, gm_stream stream) {
  if (!ctx || n_slab < 0 || n_r < 0) return gm_fail(GM_EINVAL, "gm_lut_points: bad arguments");
  if (n_slab == 0 || n_r == 0) return GM_OK;
  if (!r_dry || !lam || !m_dry || !gf || !x || !eps)
    return gm_fail(GM_EINVAL, "gm_lut_points: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  const int64_t n = n_slab * n_r;
  hipLaunchKernelGGL(gm::k_lut_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n_slab, n_r, r_dry, lam, m_dry, gf, mw_re, mw_im, x,
                     eps);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "k_lut_points: %s", hipGetErrorString(e));
  return GM_OK;
}

int gm_lut_bulk(gm_ctx* ctx, int64_t n_slab, int n_r, int n_ang, int n_psd, const double* w,
                int w_per_slab, const double* x, const double* lam, const double* props,
                const double* p11, const double* ang_deg_host, int ng, int n_mom, float* out,
                gm_stream stream) {
  if (!ctx || n_slab < 0 || n_r < 1 || n_ang < 2 || n_psd < 1 || !ang_deg_host)
    return gm_fail(GM_EINVAL, "gm_lut_bulk: bad arguments");
  if (n_mom < 1 || n_mom > gm::kLutMaxMom || n_mom > ng)
    return gm_fail(GM_EINVAL, "gm_lut_bulk: n_mom must be in [1, min(ng, %d)]", gm::kLutMaxMom);
  if (n_slab == 0) return GM_OK;
  if (!w || !x || !lam || !props || !p11 || !out)
    return gm_fail(GM_EINVAL, "gm_lut_bulk: NULL array");
  for (int a = 1; a < n_ang; ++a)
    if (!(ang_deg_host[a] > ang_deg_host[a - 1]))
      return gm_fail(GM_EINVAL, "gm_lut_bulk: angles must increase");
  const int64_t ncb = (n_ang + gm::kLutFeat + 63) / 64;
  const int64_t n_ent = n_slab * n_psd;
  if (n_slab * ncb > 0x7fffffffLL || n_ent > 0x7fffffffLL)
    return gm_fail(GM_EINVAL, "gm_lut_bulk: too many slabs for one call");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  int rc = gm::gsf_tables(ctx, ng, n_ang, ang_deg_host, s);
  if (rc) return rc;
  rc = ctx->reserve(ctx->lut_tmp, sizeof(double) * (size_t)n_ent * (n_ang + gm::kLutFeat));
  if (rc) return rc;
  double* sums = (double*)ctx->lut_tmp.ptr;
  size_t slot;
  ctx->mark_begin(gm::kKLut, s, &slot);
  hipLaunchKernelGGL(gm::k_lut_contract,
                     dim3((unsigned)(n_slab * ncb),
                          (unsigned)((n_psd + 16 * gm::kLutRowTiles - 1) / (16 * gm::kLutRowTiles))),
                     dim3(64), 0, s, n_slab, n_r, n_ang, n_psd, w,
                     w_per_slab ? (int64_t)n_psd * n_r : (int64_t)0, x, lam, props, p11, sums);
  hipLaunchKernelGGL(gm::k_lut_finish, dim3((unsigned)n_ent), dim3(64), 0, s, n_ang, ng, n_mom,
                     (const double*)ctx->gsf_tab.ptr, sums, out);
  ctx->mark_end(slot, s);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "gm_lut_bulk: %s", hipGetErrorString(e));
  return GM_OK;
}

}  // extern "C"
