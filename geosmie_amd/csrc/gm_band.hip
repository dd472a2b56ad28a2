// Band averaging of monochromatic table variables on gfx950.
//
// bandaverage.doAverage (src/geosmie/bandaverage.py:18-50) replaces a
// variable's spectrum by its mean over a band: n_sub equally spaced
// wavenumbers across the band (clipped to the table's range), the spectrum
// linearly interpolated to each (scipy interp1d, which for 1-D float64 data
// is numpy.interp), then summed and divided by n_sub.  The reference loops
// over (rh, band, bin, variable) in Python; here one thread computes one
// (series, band) mean, a series being one variable of one (bin, rh).
//
// Bit-for-bit with numpy: the interpolation is numpy.interp's rule (exact
// node -> node value; last node -> last value; slope * (x - x_j) + y_j with
// no fused multiply-add; the NaN retry from the right node), and the sum is
// numpy's pairwise sum for n <= 128 (eight running sums seeded with the first
// eight values, combined as ((s0+s1)+(s2+s3))+((s4+s5)+(s6+s7)), then the
// remainder in order).
#include <math.h>

#include "gm_ctx.h"

namespace gm {

constexpr int kBandMaxSub = 128;

// numpy.interp for one x against ascending xp[0..n).  This file is compiled
// with -ffp-contract=off (geosmie_amd/build.py): numpy rounds the product and
// the sum separately, and a fused multiply-add would not be bit-identical.
__device__ __forceinline__ double np_interp(double x, const double* __restrict__ xp,
                                            const double* __restrict__ fp, int n) {
  if (x != x) return x;
  if (x < xp[0]) return fp[0];
  if (x > xp[n - 1]) return fp[n - 1];
  // j = last index with xp[j] <= x
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (xp[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  const int j = lo;
  if (j == n - 1) return fp[j];
  if (xp[j] == x) return fp[j];
  const double slope = (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j]);
  double r = slope * (x - xp[j]) + fp[j];
  if (r != r) {
    r = slope * (x - xp[j + 1]) + fp[j + 1];
    if (r != r && fp[j] == fp[j + 1]) r = fp[j];
  }
  return r;
}

__global__ __launch_bounds__(256) void k_band_average(int64_t n_series, int n_lam,
                                                      const double* __restrict__ xp,
                                                      const double* __restrict__ y, int n_band,
                                                      int n_sub,
                                                      const double* __restrict__ samples,
                                                      double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_series * n_band) return;
  const int64_t s = t / n_band;
  const int b = (int)(t - s * n_band);
  const double* fp = y + s * n_lam;
  const double* xs = samples + (int64_t)b * n_sub;
  double res;
  if (n_sub < 8) {
    res = 0.0;
    for (int i = 0; i < n_sub; ++i) res += np_interp(xs[i], xp, fp, n_lam);
  } else {
    double r[8];
    for (int k = 0; k < 8; ++k) r[k] = np_interp(xs[k], xp, fp, n_lam);
    const int whole = n_sub - n_sub % 8;
    int i = 8;
    for (; i < whole; i += 8)
      for (int k = 0; k < 8; ++k) r[k] += np_interp(xs[i + k], xp, fp, n_lam);
    res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n_sub; ++i) res += np_interp(xs[i], xp, fp, n_lam);
  }
  out[t] = res / (double)n_sub;
}

}  // namespace gm

extern "C" {

int gm_band_average(gm_ctx* ctx, int64_t n_series, int n_lam, const double* wavenumber,
                    const double* y, int n_band, int n_sub, const double* samples, double* out,
                    gm_stream stream) {
  if (!ctx || n_series < 0 || n_lam < 1 || n_band < 0 || n_sub < 1 || n_sub > gm::kBandMaxSub)
    return gm_fail(-1, "gm_band_average: bad arguments (n_sub must be 1..%d)", gm::kBandMaxSub);
  if (n_series == 0 || n_band == 0) return 0;
  if (!wavenumber || !y || !samples || !out) return gm_fail(-1, "gm_band_average: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  const int64_t n = n_series * (int64_t)n_band;
  if (n > ((int64_t)1 << 38)) return gm_fail(-1, "gm_band_average: too many (series, band) pairs");
  hipLaunchKernelGGL(gm::k_band_average, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n_series, n_lam, wavenumber, y, n_band, n_sub, samples,
                     out);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "k_band_average: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
