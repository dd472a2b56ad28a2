// Angle functions pi_n, tau_n and efficiencies from given coefficients.
#include "gm_complex.h"
#include "gm_ctx.h"

namespace gm {

// pi'_n(u), tau'_n(u) for n = 1..nmax_cap at every angle, premultiplied by
// (2n+1)/(n(n+1)) (src/pymiecoated/pymiecoated/mie_props.py:166-231):
//   pi_1 = 1, pi_2 = 3u, pi_k = (2k-1)/(k-1) u pi_{k-1} - k/(k-1) pi_{k-2}
//   tau_1 = u, tau_2 = 6u^2 - 3, tau_k = k u pi_k - (k+1) pi_{k-1}
// The table is shared by every size (the reference's per-nmax tables,
// mie_coated.py:160-179, are its prefixes).  Layout [n-1][angle]: one K-row
// of the S1/S2 contraction is contiguous over angles.  One thread per angle.
__global__ void k_angle_tables(int n_ang, const double* __restrict__ cost, int nmax_cap,
                               double* __restrict__ pi_t, double* __restrict__ tau_t) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_ang) return;
  const double u = cost[j];
  double p_pp = 0.0, p_p = 0.0;
  for (int k = 1; k <= nmax_cap; ++k) {
    double p, t;
    if (k == 1) {
      p = 1.0;
      t = u;
    } else if (k == 2) {
      p = 3.0 * u;
      t = 6.0 * u * u - 3.0;
    } else {
      double nf = (double)(k - 1);  // the reference's loop variable n (index k-1)
      p = (2.0 * nf + 1.0) / nf * p_p * u - (nf + 1.0) / nf * p_pp;
      t = (nf + 1.0) * u * p - (nf + 2.0) * p_p;
    }
    double w = (double)(2 * k + 1) / (double)(k * (k + 1));
    pi_t[(size_t)(k - 1) * n_ang + j] = p * w;
    tau_t[(size_t)(k - 1) * n_ang + j] = t * w;
    p_pp = p_p;
    p_p = p;
  }
}

// Efficiencies of given coefficient sets (mie_props.py:28-70), one lane per set.
__global__ void k_props_from_coeffs(int64_t n, const int64_t* __restrict__ off,
                                    const double* __restrict__ an, const double* __restrict__ bn,
                                    const double* __restrict__ size, double* __restrict__ props) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int64_t o = off[p];
  int nmax = (int)(off[p + 1] - o);
  double sext = 0, ssca = 0, asy1 = 0, asy2 = 0;
  cd qb = cd{0, 0}, ap = cd{0, 0}, bp = cd{0, 0};
  for (int k = 1; k <= nmax; ++k) {
    cd a = cd{an[2 * (o + k - 1)], an[2 * (o + k - 1) + 1]};
    cd b = cd{bn[2 * (o + k - 1)], bn[2 * (o + k - 1) + 1]};
    double dn = (double)k, cn = 2.0 * dn + 1.0;
    sext = fma(cn, a.re + b.re, sext);
    ssca = fma(cn, norm2(a) + norm2(b), ssca);
    qb = qb + (a - b) * ((k & 1) ? -cn : cn);
    if (k > 1) {
      double dm = dn - 1.0;
      asy1 = fma(dm * (dm + 2.0) / (dm + 1.0),
                 ap.re * a.re + ap.im * a.im + bp.re * b.re + bp.im * b.im, asy1);
    }
    asy2 = fma(cn / (dn * (dn + 1.0)), a.re * b.re + a.im * b.im, asy2);
    ap = a;
    bp = b;
  }
  double y2 = size[p] * size[p];
  double qext = 2.0 * sext / y2, qsca = 2.0 * ssca / y2;
  double qbv = (qb.re * qb.re + qb.im * qb.im) / y2;
  props[0 * n + p] = qext;
  props[1 * n + p] = qsca;
  props[2 * n + p] = qext - qsca;
  props[3 * n + p] = qbv;
  props[4 * n + p] = 4.0 / y2 * (asy1 + asy2) / qsca;
  props[5 * n + p] = qbv / qsca;
}

}  // namespace gm

extern "C" {

int gm_angle_tables(gm_ctx* ctx, int n_ang, const double* cos_theta, int nmax_cap, double* pi_t,
                    double* tau_t, gm_stream stream) {
  if (!ctx || n_ang < 0 || nmax_cap < 0 || (n_ang > 0 && (!cos_theta || !pi_t || !tau_t)))
    return gm_fail(-1, "gm_angle_tables: bad arguments");
  if (n_ang == 0 || nmax_cap == 0) return 0;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(gm::k_angle_tables, dim3((n_ang + 63) / 64), dim3(64), 0,
                     (hipStream_t)stream, n_ang, cos_theta, nmax_cap, pi_t, tau_t);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "k_angle_tables: %s", hipGetErrorString(e));
  return 0;
}

int gm_mie_props(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const double* an,
                 const double* bn, const double* size, double* props, gm_stream stream) {
  if (!ctx || n < 0 || (n > 0 && (!coef_off || !an || !bn || !size || !props)))
    return gm_fail(-1, "gm_mie_props: bad arguments");
  if (n == 0) return 0;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(gm::k_props_from_coeffs, dim3((unsigned)((n + 63) / 64)), dim3(64), 0,
                     (hipStream_t)stream, n, coef_off, an, bn, size, props);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "k_props_from_coeffs: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
