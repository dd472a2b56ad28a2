// Size-distribution integration (runoptics integratePSD) and phase-matrix
// normalisation on gfx950.
//
// Every integral of src/geosmie/dointegration.py:1064-1209 is a weighted sum
// over the size grid, so a batch of PSDs against per-size quantities is one
// contraction  out[k][c] = sum_p w[k][p] v[p][c]:
//   - v = Mueller planes [p][4 * n_ang] for the phase-matrix integrals
//     (:1164-1166);
//   - v = 11 per-size features (k_scalar_features) for the efficiency, g,
//     cross-section and radius integrals (:1102-1190), finished by
//     k_scalar_finish with the reference's weighting rules.
// Large PSD batches (LUT sweeps) use FP64 MFMA 16x16x4 tiles; one or a few
// PSDs (a runoptics slab) use a split-K VALU kernel streaming v at HBM rate.
// Both reduce in a fixed order (bitwise reproducible).
#include <algorithm>

#include "gm_complex.h"
#include "gm_ctx.h"

typedef double double4_ __attribute__((ext_vector_type(4)));

namespace gm {

constexpr int kSmallPsd = 8;     // VALU path up to this many PSDs

// out_part[split][k][c] = sum_{p in chunk} w[k][p] v[p][c]   (k < n_psd <= kSmallPsd)
__global__ __launch_bounds__(256) void k_contract_small(int n_psd, int64_t n_pts, int n_cols,
                                                        int chunk,
                                                        const double* __restrict__ w,
                                                        const double* __restrict__ v,
                                                        double* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.y * chunk;
  const int64_t p1 = min(p0 + (int64_t)chunk, n_pts);
  if (c >= n_cols) return;
  double acc[kSmallPsd];
#pragma unroll
  for (int k = 0; k < kSmallPsd; ++k) acc[k] = 0.0;
  for (int64_t p = p0; p < p1; ++p) {
    double x = v[(size_t)p * n_cols + c];
#pragma unroll
    for (int k = 0; k < kSmallPsd; ++k)
      if (k < n_psd) acc[k] = fma(w[(size_t)k * n_pts + p], x, acc[k]);
  }
  for (int k = 0; k < n_psd; ++k)
    part[((size_t)blockIdx.y * n_psd + k) * n_cols + c] = acc[k];
}

__global__ void k_reduce_splits(int n_split, int64_t n_out, const double* __restrict__ part,
                                double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  double s = 0.0;
  for (int q = 0; q < n_split; ++q) s += part[(size_t)q * n_out + i];
  out[i] = s;
}

// Many-split variant: one wavefront per output, lane l adds splits l, l + 64,
// ... in order, then the 64 lane sums combine in a fixed butterfly order
// (deterministic; used when there are too many splits for one serial pass).
__global__ __launch_bounds__(64) void k_reduce_splits_wide(int n_split, int64_t n_out,
                                                           const double* __restrict__ part,
                                                           double* __restrict__ out) {
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  double s = 0.0;
  for (int q = lane; q < n_split; q += 64) s += part[(size_t)q * n_out + i];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
  if (lane == 0) out[i] = s;
}

// Segmented contraction: row r integrates only its own points
// [start_r, start_r + count_r):  out[r][c] = sum_p w[r][p] v[p][c].  Each
// row's range is cut into chunks; block (column block, chunk q) writes the
// chunk's partial sum, then k_reduce_segs adds a row's chunks in order.
constexpr int kMaxSegs = 64;
struct SegArgs {
  int n;                          // rows
  int chunk;                      // points per chunk
  int64_t start[kMaxSegs], count[kMaxSegs];
  int first[kMaxSegs + 1];        // first chunk of each row; first[n] = total
};

__global__ __launch_bounds__(256) void k_contract_segs(SegArgs a, int64_t n_pts, int n_cols,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  int r = 0;
  while (r + 1 < a.n && a.first[r + 1] <= q) ++r;
  const int64_t p0 = a.start[r] + (int64_t)(q - a.first[r]) * a.chunk;
  const int64_t p1 = min(p0 + (int64_t)a.chunk, a.start[r] + a.count[r]);
  if (c >= n_cols) return;
  const double* wr = w + (size_t)r * n_pts;
  double acc = 0.0;
  for (int64_t p = p0; p < p1; ++p) acc = fma(wr[p], v[(size_t)p * n_cols + c], acc);
  part[(size_t)q * n_cols + c] = acc;
}

__global__ void k_reduce_segs(SegArgs a, int n_cols, const double* __restrict__ part,
                              double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= n_cols || r >= a.n) return;
  double s = 0.0;
  for (int q = a.first[r]; q < a.first[r + 1]; ++q) s += part[(size_t)q * n_cols + c];
  out[(size_t)r * n_cols + c] = s;
}

// Batched PSDs on FP64 MFMA: a wavefront owns 16 PSDs x 64 columns.
// A = w (k rows, p = K), B = v (p = K, c columns); D[g + 4r][j] in lane j + 16g.
// Split-K: blockIdx.z covers points [z * kSplitChunkMfma, ...); partial sums
// to part[z][k][c], reduced in order by k_reduce_splits.
constexpr int kSplitChunkMfma = 2048;
__global__ __launch_bounds__(256) void k_contract_mfma(int n_psd, int64_t n_pts, int n_cols,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k0 = blockIdx.y * 16;
  const int c0 = (blockIdx.x * 4 + wave) * 64;
  if (c0 >= n_cols) return;
  const int i = lane & 15, kq = lane >> 4;
  const int ka = k0 + i;
  const int64_t pbeg = (int64_t)blockIdx.z * kSplitChunkMfma;
  const int64_t pend = min(pbeg + (int64_t)kSplitChunkMfma, n_pts);
  double* out = part + (size_t)blockIdx.z * n_psd * n_cols;
  double4_ acc[4];
  for (int t = 0; t < 4; ++t) acc[t] = (double4_){0.0, 0.0, 0.0, 0.0};
  for (int64_t p0 = pbeg; p0 < pend; p0 += 4) {
    const int64_t p = p0 + kq;
    const bool pin = p < pend;
    double a = (pin && ka < n_psd) ? w[(size_t)ka * n_pts + p] : 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = c0 + t * 16 + i;
      double b = (pin && c < n_cols) ? v[(size_t)p * n_cols + c] : 0.0;
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int c = c0 + t * 16 + i;
    if (c >= n_cols) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int k = k0 + kq + 4 * r;
      if (k < n_psd) out[(size_t)k * n_cols + c] = acc[t][r];
    }
  }
}

// Narrow variant (n_cols <= 16: the integratePSD features): one wavefront per
// (16 PSDs, chunk of kNarrowChunk points), so a batch of 10^4..10^5 points
// spreads over hundreds of wavefronts; the loads of kNarrowAhead MFMA steps
// are issued before their products, hiding their latency inside the wave.
constexpr int kNarrowChunk = 256;
constexpr int kNarrowAhead = 8;
__global__ __launch_bounds__(64) void k_contract_mfma_narrow(int n_psd, int64_t n_pts, int n_cols,
                                                             const double* __restrict__ w,
                                                             const double* __restrict__ v,
                                                             double* __restrict__ part) {
  const int lane = threadIdx.x;
  const int k0 = blockIdx.x * 16;
  const int i = lane & 15, kq = lane >> 4;
  const int ka = k0 + i;
  const int64_t pbeg = (int64_t)blockIdx.y * kNarrowChunk;
  const int64_t pend = min(pbeg + (int64_t)kNarrowChunk, n_pts);
  double4_ acc = (double4_){0.0, 0.0, 0.0, 0.0};
  for (int64_t p0 = pbeg; p0 < pend; p0 += 4 * kNarrowAhead) {
    double a[kNarrowAhead], b[kNarrowAhead];
#pragma unroll
    for (int u = 0; u < kNarrowAhead; ++u) {
      const int64_t p = p0 + 4 * u + kq;
      const bool pin = p < pend;
      a[u] = (pin && ka < n_psd) ? w[(size_t)ka * n_pts + p] : 0.0;
      b[u] = (pin && i < n_cols) ? v[(size_t)p * n_cols + i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kNarrowAhead; ++u)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
  }
  double* out = part + (size_t)blockIdx.y * n_psd * n_cols;
  if (i < n_cols) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = k0 + kq + 4 * r;
      if (k < n_psd) out[(size_t)k * n_cols + i] = acc[r];
    }
  }
}

// Per-size features for integratePSD (r = x lam / 2 pi; props planes
// qext, qsca, qabs, qb, asy, qratio):
//   0 1   1 r^2   2 r^3   3 r^4   4 qext r^2   5 qsca r^2   6 qabs r^2
//   7 qb r^2   8 g qsca r^2   9 csca qsca r^2   10 cext qsca r^2
// (csca = qsca pi r^2, cext = qext pi r^2, dointegration.py:1219-1220)
constexpr int kNFeat = 13;
// lam_pts (optional): a wavelength per point, for batches of several rows.
__global__ void k_scalar_features(int64_t n, const double* __restrict__ xx, double lam,
                                  const double* __restrict__ lam_pts,
                                  const double* __restrict__ props, double* __restrict__ v) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double pi = 3.14159265358979323846;
  double r = xx[p] * (lam_pts ? lam_pts[p] : lam) / (2.0 * pi);
  double r2 = r * r;
  double qext = props[p], qsca = props[n + p], qabs = props[2 * n + p], qb = props[3 * n + p];
  double g = props[4 * n + p];
  double* o = v + (size_t)p * kNFeat;
  o[0] = 1.0;
  o[1] = r2;
  o[2] = r2 * r;
  o[3] = r2 * r2;
  o[4] = qext * r2;
  o[5] = qsca * r2;
  o[6] = qabs * r2;
  o[7] = qb * r2;
  o[8] = g * qsca * r2;
  o[9] = (qsca * pi * r2) * qsca * r2;
  o[10] = (qext * pi * r2) * qsca * r2;
  // csca, cext area-weighted only (the weighting without the aliased Qsca)
  o[11] = (qsca * pi * r2) * r2;
  o[12] = (qext * pi * r2) * r2;
}

// Finish integratePSD for n_modes PSDs (sums[k][feature]) and mix them with
// fracs (dointegration.py:1094-1200, mixing 'dumix').  One thread, modes in
// order.  out (kNOut): num, area, volume, mass, rEff, rMass, qext, qsca, qabs,
// qb, g, csca, cext, bsca, bext, bbck.
constexpr int kNOut = 16;
// quirks bit 0 (GM_PSD_QUIRK_ALIAS): csca and cext carry the Qsca weight they
// inherit from the aliased thisweight; without it they are area-weighted like
// the other efficiencies.
__device__ void finish_modes(int n_modes, const double* __restrict__ sums,
                             const double* __restrict__ fracs, const double* __restrict__ reff0,
                             double rhop0, double rhop, int quirks, double* __restrict__ out) {
  const double pi = 3.14159265358979323846;
  double acc[kNOut];
  for (int q = 0; q < kNOut; ++q) acc[q] = 0.0;
  for (int k = 0; k < n_modes; ++k) {
    const double* s = sums + (size_t)k * kNFeat;
    double num = s[0], S2 = s[1], S3 = s[2], S4 = s[3];
    double th[kNOut];
    th[0] = num;
    th[1] = pi * S2 / num;                    // area (:1107)
    th[2] = 4.0 / 3.0 * pi * S3 / num;        // volume (:1108)
    th[3] = th[2] * rhop;                     // mass (:1109)
    th[4] = S3 / S2;                          // rEff (:1114)
    double reff_mass = S4 / S3;               // (:1122)
    th[5] = 4.0 / 3.0 * pi * rhop * reff_mass * reff_mass * reff_mass;   // rMass
    double rMass0 = 4.0 / 3.0 * pi * rhop0 * reff0[k] * reff0[k] * reff0[k];
    // efficiencies, area weighted then normalised by the weight sum (:1188-1190)
    th[6] = s[4] / S2;                        // qext
    th[7] = s[5] / S2;                        // qsca
    th[8] = s[6] / S2;                        // qabs
    // qb from the backed-out P11 at 180 deg (:1133-1145, 1168-1169)
    th[9] = (4.0 * pi * s[7] / s[5]) * th[7] / (4.0 * pi);
    // g and every key after it carry the extra Qsca weight (:1111-1112, 1154-1157)
    th[10] = s[8] / s[5];                     // g
    if (quirks & 1) {
      th[11] = s[9] / s[5];                   // csca
      th[12] = s[10] / s[5];                  // cext
    } else {
      th[11] = s[11] / S2;
      th[12] = s[12] / S2;
    }
    double mc = 1.0 / rhop / th[4] * th[5] / rMass0;   // (:1193)
    th[13] = 0.75 * th[7] * mc;               // bsca
    th[14] = 0.75 * th[6] * mc;               // bext
    th[15] = 0.75 * th[9] * mc / (4.0 * pi);  // bbck
    for (int q = 0; q < kNOut; ++q) acc[q] += (fracs ? fracs[k] : 1.0) * th[q];
  }
  for (int q = 0; q < kNOut; ++q) out[q] = acc[q];
}

__global__ void k_scalar_finish(int n_modes, const double* __restrict__ sums,
                                const double* __restrict__ fracs,
                                const double* __restrict__ reff0, double rhop0, double rhop,
                                int quirks, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  finish_modes(n_modes, sums, fracs, reff0, rhop0, rhop, quirks, out);
}

// One row per thread: rows are independent single-mode integrations.
// rhop_rows (or nullptr: rhop for every row) is each row's wet density.
__global__ void k_scalar_finish_rows(int64_t n_rows, const double* __restrict__ sums,
                                     const double* __restrict__ reff0, double rhop0, double rhop,
                                     const double* __restrict__ rhop_rows, int quirks,
                                     double* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  finish_modes(1, sums + r * kNFeat, nullptr, reff0 + r, rhop0, rhop_rows ? rhop_rows[r] : rhop,
               quirks, out + r * kNOut);
}

// Phase-matrix normalisation, dointegration.py:978-988, per entry:
// p11n = 2 p11 / trapz(p11 sin(theta), theta); other elements * p11n / p11.
// pmat[e][4][n_ang] (s11, s12, s33, s34) in place; pback[e][6] in the table's
// element order p11, p12, p33, p34, p22, p44 at the last angle.
__global__ __launch_bounds__(256) void k_pmat_normalize(int64_t n_ent, int n_ang,
                                                        const double* __restrict__ theta,
                                                        double* __restrict__ pmat,
                                                        double* __restrict__ pback) {
  __shared__ double red[256];
  const int64_t e = blockIdx.x;
  if (e >= n_ent) return;
  double* P = pmat + (size_t)e * 4 * n_ang;
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
    double p11 = P[j];
    double p11n = 2.0 * p11 / integ;
    double f = p11n / p11;
    P[n_ang + j] *= f;
    P[2 * n_ang + j] *= f;
    P[3 * n_ang + j] *= f;
    P[j] = p11n;
  }
  __syncthreads();
  if (threadIdx.x == 0 && pback) {
    int j = n_ang - 1;
    double* b = pback + (size_t)e * 6;
    b[0] = P[j];
    b[1] = P[n_ang + j];
    b[2] = P[2 * n_ang + j];
    b[3] = P[3 * n_ang + j];
    b[4] = P[j];
    b[5] = P[2 * n_ang + j];
  }
}

}  // namespace gm

extern "C" {

int gm_psd_contract(gm_ctx* ctx, int n_psd, int64_t n_pts, int n_cols, const double* w,
                    const double* v, double* out, gm_stream stream) {
  if (!ctx || n_psd < 0 || n_pts < 0 || n_cols < 0)
    return gm_fail(-1, "gm_psd_contract: bad arguments");
  if (n_psd == 0 || n_cols == 0) return 0;
  if (!w || !v || !out) return gm_fail(-1, "gm_psd_contract: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  size_t slot;
  ctx->mark_begin(gm::kKPsd, s, &slot);
  if (n_psd <= gm::kSmallPsd) {
    const int chunk = ctx->tune_psd_chunk;
    int nsplit = (int)std::max<int64_t>(1, (n_pts + chunk - 1) / chunk);
    size_t bytes = sizeof(double) * (size_t)nsplit * n_psd * n_cols;
    int rc = ctx->reserve(ctx->gemm_tmp, bytes);
    if (rc) return rc;
    double* part = (double*)ctx->gemm_tmp.ptr;
    hipLaunchKernelGGL(gm::k_contract_small, dim3((n_cols + 255) / 256, nsplit), dim3(256), 0, s,
                       n_psd, n_pts, n_cols, chunk, w, v, part);
    int64_t n_out = (int64_t)n_psd * n_cols;
    hipLaunchKernelGGL(gm::k_reduce_splits, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0,
                       s, nsplit, n_out, part, out);
  } else if (n_cols <= 16) {
    const int64_t nsplit64 = (n_pts + gm::kNarrowChunk - 1) / gm::kNarrowChunk;
    if (nsplit64 > 65535) return gm_fail(-1, "gm_psd_contract: too many points for one call");
    const int nsplit = (int)std::max<int64_t>(1, nsplit64);
    size_t bytes = sizeof(double) * (size_t)nsplit * n_psd * n_cols;
    int rc = ctx->reserve(ctx->gemm_tmp, bytes);
    if (rc) return rc;
    double* part = (double*)ctx->gemm_tmp.ptr;
    hipLaunchKernelGGL(gm::k_contract_mfma_narrow, dim3((n_psd + 15) / 16, nsplit), dim3(64), 0,
                       s, n_psd, n_pts, n_cols, w, v, part);
    int64_t n_out = (int64_t)n_psd * n_cols;
    if (nsplit > 64)
      hipLaunchKernelGGL(gm::k_reduce_splits_wide, dim3((unsigned)n_out), dim3(64), 0, s, nsplit,
                         n_out, part, out);
    else
      hipLaunchKernelGGL(gm::k_reduce_splits, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0,
                         s, nsplit, n_out, part, out);
  } else {
    int nsplit =
        (int)std::max<int64_t>(1, (n_pts + gm::kSplitChunkMfma - 1) / gm::kSplitChunkMfma);
    size_t bytes = sizeof(double) * (size_t)nsplit * n_psd * n_cols;
    int rc = ctx->reserve(ctx->gemm_tmp, bytes);
    if (rc) return rc;
    double* part = (double*)ctx->gemm_tmp.ptr;
    dim3 grid((n_cols + 255) / 256, (n_psd + 15) / 16, nsplit);
    hipLaunchKernelGGL(gm::k_contract_mfma, grid, dim3(256), 0, s, n_psd, n_pts, n_cols, w, v,
                       part);
    int64_t n_out = (int64_t)n_psd * n_cols;
    hipLaunchKernelGGL(gm::k_reduce_splits, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0,
                       s, nsplit, n_out, part, out);
  }
  ctx->mark_end(slot, s);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "gm_psd_contract: %s", hipGetErrorString(e));
  return 0;
}

int gm_integrate_psd(gm_ctx* ctx, int64_t n_pts, int n_modes, const double* xx, const double* props,
                     const double* psd, const double* fracs, const double* reff0, double lam,
                     double rhop0, double rhop, double* out, gm_stream stream) {
  if (!ctx || n_pts <= 0 || n_modes <= 0 || !xx || !props || !psd || !fracs || !reff0 || !out)
    return gm_fail(-1, "gm_integrate_psd: bad arguments");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  size_t feat_bytes = sizeof(double) * (size_t)n_pts * gm::kNFeat;
  size_t sum_bytes = sizeof(double) * (size_t)n_modes * gm::kNFeat;
  int rc = ctx->reserve(ctx->feat_tmp, feat_bytes + sum_bytes);
  if (rc) return rc;
  double* v = (double*)ctx->feat_tmp.ptr;
  double* sums = v + (size_t)n_pts * gm::kNFeat;
  hipLaunchKernelGGL(gm::k_scalar_features, dim3((unsigned)((n_pts + 255) / 256)), dim3(256), 0, s,
                     n_pts, xx, lam, (const double*)nullptr, props, v);
  rc = gm_psd_contract(ctx, n_modes, n_pts, gm::kNFeat, psd, v, sums, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(gm::k_scalar_finish, dim3(1), dim3(64), 0, s, n_modes, sums, fracs, reff0,
                     rhop0, rhop, ctx->psd_quirks, out);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "gm_integrate_psd: %s", hipGetErrorString(e));
  return 0;
}

int gm_pmatrix_normalize(gm_ctx* ctx, int64_t n_entries, int n_ang, const double* theta_rad,
                         double* pmat, double* pback, gm_stream stream) {
  if (!ctx || n_entries < 0 || n_ang < 2 || !theta_rad || !pmat)
    return gm_fail(-1, "gm_pmatrix_normalize: bad arguments");
  if (n_entries == 0) return 0;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(gm::k_pmat_normalize, dim3((unsigned)n_entries), dim3(256), 0,
                     (hipStream_t)stream, n_entries, n_ang, theta_rad, pmat, pback);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "gm_pmatrix_normalize: %s", hipGetErrorString(e));
  return 0;
}

int gm_integrate_psd_rows(gm_ctx* ctx, int64_t n_rows, int64_t n_pts, const double* xx,
                          const double* lam_pts, const double* props, const double* psd,
                          const double* reff0, double rhop0, double rhop,
                          const double* rhop_rows, double* out, gm_stream stream) {
  if (!ctx || n_pts <= 0 || n_rows <= 0 || n_rows > 0x7fffffff || !xx || !lam_pts || !props ||
      !psd || !reff0 || !out)
    return gm_fail(-1, "gm_integrate_psd_rows: bad arguments");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  size_t feat_bytes = sizeof(double) * (size_t)n_pts * gm::kNFeat;
  size_t sum_bytes = sizeof(double) * (size_t)n_rows * gm::kNFeat;
  int rc = ctx->reserve(ctx->feat_tmp, feat_bytes + sum_bytes);
  if (rc) return rc;
  double* v = (double*)ctx->feat_tmp.ptr;
  double* sums = v + (size_t)n_pts * gm::kNFeat;
  hipLaunchKernelGGL(gm::k_scalar_features, dim3((unsigned)((n_pts + 255) / 256)), dim3(256), 0, s,
                     n_pts, xx, 0.0, lam_pts, props, v);
  rc = gm_psd_contract(ctx, (int)n_rows, n_pts, gm::kNFeat, psd, v, sums, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(gm::k_scalar_finish_rows, dim3((unsigned)((n_rows + 63) / 64)), dim3(64), 0, s,
                     n_rows, sums, reff0, rhop0, rhop, rhop_rows, ctx->psd_quirks, out);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "gm_integrate_psd_rows: %s", hipGetErrorString(e));
  return 0;
}

int gm_psd_contract_rows(gm_ctx* ctx, int n_rows, const int64_t* start_host,
                         const int64_t* count_host, int64_t n_pts, int n_cols, const double* w,
                         const double* v, double* out, gm_stream stream) {
  if (!ctx || n_rows < 0 || n_rows > gm::kMaxSegs || n_pts < 0 || n_cols < 0)
    return gm_fail(-1, "gm_psd_contract_rows: bad arguments (at most %d rows)", gm::kMaxSegs);
  if (n_rows == 0 || n_cols == 0) return 0;
  if (!start_host || !count_host || !w || !v || !out)
    return gm_fail(-1, "gm_psd_contract_rows: NULL array");
  gm::SegArgs a;
  a.n = n_rows;
  a.chunk = ctx->tune_psd_chunk;
  a.first[0] = 0;
  for (int r = 0; r < n_rows; ++r) {
    if (start_host[r] < 0 || count_host[r] < 1 || start_host[r] + count_host[r] > n_pts)
      return gm_fail(-1, "gm_psd_contract_rows: row %d range outside [0, n_pts)", r);
    a.start[r] = start_host[r];
    a.count[r] = count_host[r];
    const int64_t nc = (count_host[r] + a.chunk - 1) / a.chunk;
    if (a.first[r] + nc > 0x7fffffffLL) return gm_fail(-1, "gm_psd_contract_rows: too many chunks");
    a.first[r + 1] = a.first[r] + (int)nc;
  }
  for (int r = n_rows; r < gm::kMaxSegs; ++r) a.start[r] = a.count[r] = 0;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  const int total = a.first[n_rows];
  int rc = ctx->reserve(ctx->gemm_tmp, sizeof(double) * (size_t)total * n_cols);
  if (rc) return rc;
  double* part = (double*)ctx->gemm_tmp.ptr;
  size_t slot;
  ctx->mark_begin(gm::kKPsd, s, &slot);
  const unsigned cb = (unsigned)((n_cols + 255) / 256);
  hipLaunchKernelGGL(gm::k_contract_segs, dim3(cb, (unsigned)total), dim3(256), 0, s, a, n_pts,
                     n_cols, w, v, part);
  hipLaunchKernelGGL(gm::k_reduce_segs, dim3(cb, (unsigned)n_rows), dim3(256), 0, s, a, n_cols,
                     (const double*)part, out);
  ctx->mark_end(slot, s);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "gm_psd_contract_rows: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
