// S1/S2 amplitude sums as a real GEMM on the gfx950 matrix cores.
//
// S1(p,j) = sum_n a_n pi'_n(j) + b_n tau'_n(j),  S2 = sum_n a_n tau'_n + b_n pi'_n
// (src/pymiecoated/pymiecoated/mie_props.py:133-149, looped over angles by
// mie_coated.py:31-39).  With U = a + b, V = a - b, P = pi' + tau',
// M = pi' - tau':  S1 + S2 = sum U P,  S1 - S2 = sum V M,  so the pair is two
// real GEMMs  [Re U; Im U](2P x K) x P(K x N_ang)  and the same for V, M, with
// K = n (ragged per point, zero-padded to the tile maximum).
//
// FP64 path: v_mfma_f64_16x16x4_f64.  A wavefront owns 8 points (16 rows =
// re/im interleaved) x 64 angles (four 16-wide sub-tiles) and walks K in steps
// of 4.  Operand layouts of v_mfma_f64_16x16x4_f64 (measured on gfx950):
// A[i][k] from lane i + 16k, B[k][j] from lane j + 16k, D[g + 4r][j] in lane
// j + 16g, element r  (the f32/bf16 16x16 variants use D[4g + r][j] instead).
// Epilogue modes: write S1, S2 (complex), or the four independent Mueller
// elements s11, s12, s33, s34 (dointegration.py:1044-1050).
#include "gm_complex.h"
#include "gm_ctx.h"

typedef double double4_ __attribute__((ext_vector_type(4)));

namespace gm {

enum { kOutS12 = 0, kOutMueller = 1 };

__global__ __launch_bounds__(256) void k_s12_f64(
    int64_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ nmax,
    const double* __restrict__ an, const double* __restrict__ bn, int n_ang,
    const double* __restrict__ pi_t, const double* __restrict__ tau_t, int nmax_cap, int mode,
    double* __restrict__ out0, double* __restrict__ out1) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t p0 = (int64_t)blockIdx.x * 8;
  const int j0 = (blockIdx.y * 4 + wave) * 64;
  if (j0 >= n_ang) return;  // whole wavefront past the last angle
  const int i = lane & 15;   // A row / B column / D column
  const int kq = lane >> 4;  // k within the 4-step, D row group
  // The f64 16x16x4 result row of (lane group g, element r) is g + 4r.  A row
  // i = g + 4r is assigned to point p0 + 2g + (r >> 1), component r & 1, so
  // lane group g ends up holding (re, im) of points p0+2g and p0+2g+1.
  const int64_t pa = p0 + 2 * (i & 3) + (i >> 3);
  const bool pa_ok = pa < n;
  const int nm_a = pa_ok ? nmax[pa] : 0;
  const int64_t oa = pa_ok ? off[pa] : 0;
  const int comp = (i >> 2) & 1;
  // tile K = max nmax of the 8 points
  int kmax = 0;
  for (int q = 0; q < 8; ++q) {
    int64_t pq = p0 + q;
    if (pq < n) kmax = max(kmax, nmax[pq]);
  }
  kmax = min(kmax, nmax_cap);
  double4_ accp[4], accm[4];
  for (int t = 0; t < 4; ++t) {
    accp[t] = (double4_){0.0, 0.0, 0.0, 0.0};
    accm[t] = (double4_){0.0, 0.0, 0.0, 0.0};
  }
  for (int k0 = 0; k0 < kmax; k0 += 4) {
    const int kk = k0 + kq;
    double ua = 0.0, va = 0.0;
    if (kk < nm_a) {
      double ar = an[2 * (oa + kk) + comp];
      double br = bn[2 * (oa + kk) + comp];
      ua = ar + br;
      va = ar - br;
    }
    const bool krow = kk < kmax;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int j = j0 + t * 16 + i;
      double pp = 0.0, tt = 0.0;
      if (krow && j < n_ang) {
        pp = pi_t[(size_t)kk * n_ang + j];
        tt = tau_t[(size_t)kk * n_ang + j];
      }
      accp[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua, pp + tt, accp[t], 0, 0, 0);
      accm[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va, pp - tt, accm[t], 0, 0, 0);
    }
  }
  // epilogue: elements (0,1) = point p0+2kq (re,im), (2,3) = point p0+2kq+1 (re,im)
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int j = j0 + t * 16 + i;
    if (j >= n_ang) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t p = p0 + 2 * kq + h;
      if (p >= n) continue;
      cd sp = cd{accp[t][2 * h], accp[t][2 * h + 1]};
      cd sm = cd{accm[t][2 * h], accm[t][2 * h + 1]};
      cd s1 = (sp + sm) * 0.5;
      cd s2 = (sp - sm) * 0.5;
      size_t o = (size_t)p * n_ang + j;
      if (mode == kOutS12) {
        out0[2 * o] = s1.re;
        out0[2 * o + 1] = s1.im;
        out1[2 * o] = s2.re;
        out1[2 * o + 1] = s2.im;
      } else {
        double a1 = norm2(s1), a2 = norm2(s2);
        cd c = s1 * conj(s2);
        size_t q = (size_t)p * 4 * n_ang + j;  // layout [point][element][angle]
        out0[q] = 0.5 * (a1 + a2);              // s11 (= s22)
        out0[q + n_ang] = 0.5 * (a2 - a1);      // s12
        out0[q + 2 * n_ang] = c.re;             // s33 (= s44)
        out0[q + 3 * n_ang] = c.im;             // s34 = -Im(conj(S1) S2)
      }
    }
  }
}

}  // namespace gm

extern "C" {

// See include/geosmie.h.
int gm_mie_s12(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
               const double* an, const double* bn, int n_ang, const double* pi_t,
               const double* tau_t, int nmax_cap, double* s1, double* s2, int precision,
               gm_stream stream) {
  if (!ctx || n < 0 || n_ang < 0) return gm_fail(-1, "gm_mie_s12: bad arguments");
  if (n == 0 || n_ang == 0) return 0;
  if (!coef_off || !nmax || !an || !bn || !pi_t || !tau_t || !s1 || !s2)
    return gm_fail(-1, "gm_mie_s12: NULL array");
  if (precision != 0) return gm_fail(-1, "gm_mie_s12: precision %d not available", precision);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  dim3 grid((unsigned)((n + 7) / 8), (unsigned)((n_ang + 255) / 256));
  size_t slot;
  ctx->mark_begin(gm::kKS12, (hipStream_t)stream, &slot);
  hipLaunchKernelGGL(gm::k_s12_f64, grid, dim3(256), 0, (hipStream_t)stream, n, coef_off, nmax,
                     an, bn, n_ang, pi_t, tau_t, nmax_cap, (int)gm::kOutS12, s1, s2);
  ctx->mark_end(slot, (hipStream_t)stream);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "k_s12_f64: %s", hipGetErrorString(e));
  return 0;
}

int gm_s12_mueller(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
                   const double* an, const double* bn, int n_ang, const double* pi_t,
                   const double* tau_t, int nmax_cap, double* mueller, int precision,
                   gm_stream stream) {
  if (!ctx || n < 0 || n_ang < 0) return gm_fail(-1, "gm_s12_mueller: bad arguments");
  if (n == 0 || n_ang == 0) return 0;
  if (!coef_off || !nmax || !an || !bn || !pi_t || !tau_t || !mueller)
    return gm_fail(-1, "gm_s12_mueller: NULL array");
  if (precision != 0) return gm_fail(-1, "gm_s12_mueller: precision %d not available", precision);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  dim3 grid((unsigned)((n + 7) / 8), (unsigned)((n_ang + 255) / 256));
  size_t slot;
  ctx->mark_begin(gm::kKS12, (hipStream_t)stream, &slot);
  hipLaunchKernelGGL(gm::k_s12_f64, grid, dim3(256), 0, (hipStream_t)stream, n, coef_off, nmax,
                     an, bn, n_ang, pi_t, tau_t, nmax_cap, (int)gm::kOutMueller, mueller,
                     (double*)nullptr);
  ctx->mark_end(slot, (hipStream_t)stream);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "k_s12_f64: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
