// GRASP kernel mode of runoptics on the GPU: shape averaging of the kernel
// tables, their bilinear interpolation in refractive index, and the
// normalisation of six independent phase-matrix elements.
//
//   integrateShapes        src/geosmie/dointegration.py:74-91
//   get_interpolated       :40-72 with bilinear_interpolation :29-38
//   P normalisation, pback :978-988 (p22 and p44 are their own elements for
//                          non-spherical kernels, unlike the Mie path)
//
// Built without floating-point contraction: the interpolation combines the
// four corners as numpy evaluates the reference's expression (each product
// rounded, sums left to right), so results match it bit for bit.
#include "gm_ctx.h"

namespace gm {

// out[j] = sum_r f[r] data[r][j], r ascending (f already normalised).
__global__ __launch_bounds__(256) void k_shape_average(int n_shape, int64_t n_elem,
                                                       const double* __restrict__ f,
                                                       const double* __restrict__ data,
                                                       double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_elem) return;
  double acc = f[0] * data[j];
  for (int r = 1; r < n_shape; ++r) acc = acc + f[r] * data[(size_t)r * n_elem + j];
  out[j] = acc;
}

// Query q: corners (i, j), (i+1, j), (i+1, j+1), (i, j+1) of table
// [n_mr][n_mi][n_elem] with weights w[q][0..3] in that order:
// out[q][k] = ((w1 f1 + w2 f2) + w4 f4) + w3 f3  (dointegration.py:33-36).
// Indices are clamped to the grid here as well, so no query can read
// outside the table.
__global__ __launch_bounds__(256) void k_ri_bilinear(int64_t n_q, int n_mr, int n_mi,
                                                     int64_t n_elem,
                                                     const double* __restrict__ table,
                                                     const int32_t* __restrict__ idx,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (k >= n_elem || q >= n_q) return;
  const int i = min(max(idx[2 * q], 0), n_mr - 2);
  const int jj = min(max(idx[2 * q + 1], 0), n_mi - 2);
  const double* t00 = table + ((size_t)i * n_mi + jj) * n_elem;
  const double* t10 = table + ((size_t)(i + 1) * n_mi + jj) * n_elem;
  const double* t11 = table + ((size_t)(i + 1) * n_mi + jj + 1) * n_elem;
  const double* t01 = table + ((size_t)i * n_mi + jj + 1) * n_elem;
  const double w1 = w[4 * q], w2 = w[4 * q + 1], w3 = w[4 * q + 2], w4 = w[4 * q + 3];
  double a = w1 * t00[k];
  double b = w2 * t10[k];
  double c = w4 * t01[k];
  double d = w3 * t11[k];
  out[(size_t)q * n_elem + k] = ((a + b) + c) + d;
}

// Six elements per entry, [e][p11, p12, p22, p33, p34, p44][angle]:
// p11n = 2 p11 / trapz(p11 sin(theta), theta), the others times p11n / p11,
// pback = the 180-degree values in the order p11, p12, p33, p34, p22, p44.
__global__ __launch_bounds__(256) void k_pmat_normalize6(int64_t n_ent, int n_ang,
                                                         const double* __restrict__ theta,
                                                         double* __restrict__ pmat,
                                                         double* __restrict__ pback) {
  __shared__ double red[256];
  const int64_t e = blockIdx.x;
  if (e >= n_ent) return;
  double* P = pmat + (size_t)e * 6 * n_ang;
  double s = 0.0;
  for (int j = threadIdx.x; j < n_ang - 1; j += blockDim.x) {
    double f0 = P[j] * sin(theta[j]), f1 = P[j + 1] * sin(theta[j + 1]);
    s += (theta[j + 1] - theta[j]) * (f0 + f1) * 0.5;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double integ = red[0];
  for (int j = threadIdx.x; j < n_ang; j += blockDim.x) {
    const double p11 = P[j];
    const double p11n = 2.0 * p11 / integ;
    const double f = p11n / p11;
    for (int el = 1; el < 6; ++el) P[(size_t)el * n_ang + j] *= f;
    P[j] = p11n;
  }
  __syncthreads();
  if (threadIdx.x == 0 && pback) {
    const int j = n_ang - 1;
    double* b = pback + (size_t)e * 6;
    b[0] = P[j];
    b[1] = P[n_ang + j];
    b[2] = P[3 * n_ang + j];
    b[3] = P[4 * n_ang + j];
    b[4] = P[2 * n_ang + j];
    b[5] = P[5 * n_ang + j];
  }
}

}  // namespace gm

#define HIP_OK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return gm_fail(GM_ERUNTIME, "%s: %s", #expr, hipGetErrorString(e_));         \
  } while (0)

extern "C" {

int gm_shape_average(gm_ctx* ctx, int n_shape, int64_t n_elem, const double* w,
                     const double* data, double* out, gm_stream stream) {
  if (!ctx || n_shape < 1 || n_elem < 0 || !w || !data || !out)
    return gm_fail(GM_EINVAL, "gm_shape_average: bad arguments");
  if (n_elem == 0) return GM_OK;
  HIP_OK(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(gm::k_shape_average, dim3((unsigned)((n_elem + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n_shape, n_elem, w, data, out);
  HIP_OK(hipGetLastError());
  return GM_OK;
}

int gm_ri_bilinear(gm_ctx* ctx, int n_mr, int n_mi, int64_t n_elem, const double* table,
                   int64_t n_q, const int32_t* idx, const double* w, double* out,
                   gm_stream stream) {
  if (!ctx || n_mr < 2 || n_mi < 2 || n_elem < 0 || n_q < 0 || n_q > 65535 || !table || !idx ||
      !w || !out)
    return gm_fail(GM_EINVAL, "gm_ri_bilinear: bad arguments (grid >= 2 x 2, <= 65535 queries)");
  if (n_elem == 0 || n_q == 0) return GM_OK;
  HIP_OK(hipSetDevice(ctx->device));
  dim3 grid((unsigned)((n_elem + 255) / 256), (unsigned)n_q);
  hipLaunchKernelGGL(gm::k_ri_bilinear, grid, dim3(256), 0, (hipStream_t)stream, n_q, n_mr, n_mi,
                     n_elem, table, idx, w, out);
  HIP_OK(hipGetLastError());
  return GM_OK;
}

int gm_pmatrix_normalize6(gm_ctx* ctx, int64_t n_entries, int n_ang, const double* theta_rad,
                          double* pmat, double* pback, gm_stream stream) {
  if (!ctx || n_entries < 0 || n_ang < 2 || !theta_rad || !pmat)
    return gm_fail(GM_EINVAL, "gm_pmatrix_normalize6: bad arguments");
  if (n_entries == 0) return GM_OK;
  HIP_OK(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(gm::k_pmat_normalize6, dim3((unsigned)n_entries), dim3(256), 0,
                     (hipStream_t)stream, n_entries, n_ang, theta_rad, pmat, pback);
  HIP_OK(hipGetLastError());
  return GM_OK;
}

}  // extern "C"
