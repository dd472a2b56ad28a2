// Generalised-spherical-function expansion of phase matrices on gfx950.
//
// Replaces the per-entry process boundary of the reference (convertncdf.py
// writes a text file and runs spher_expan.x, convertncdf.py:173-189) with one
// batched device call.  spher_expan.f computes, for NG Gauss nodes x_i
// (GAUSS, :520-579) and F linearly interpolated in angle to acos(x_i)
// (LINTERPOL, :593-623; one_calc :157-168):
//   A_t[l] = sum_i w_i F_t(acos x_i) P^t_l(x_i)     (SPHER_EXPAN :313-334)
// with P^00 for F11, F44; P^22 for F22+F33; P^2-2 for F22-F33; P^02 for F12,
// F34 (GENER :363-407), then alpha/beta by the (l + 1/2) factors (:335-347)
// and normalisation by 1/alpha1(0) (:95-103).
//
// Interpolation is linear in F, so it folds into per-angle tables
//   G_t[a][l] = sum_i w_i P^t_l(x_i) c_{i,a}
// (c_{i,a} = interpolation weight of input angle a at node i), built once on
// the host per (NG, angle grid) and cached in the context.  Each entry is
// then a contraction of its 6 x n_ang input against the four tables.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gm_complex.h"
#include "gm_ctx.h"

namespace gm {

// Gauss-Legendre nodes (ascending) and weights, Newton iteration with the
// starting values of spher_expan.f:529-559.
static void gauss_nodes(int n, std::vector<double>& z, std::vector<double>& w) {
  z.assign(n, 0.0);
  w.assign(n, 0.0);
  const int ind = n % 2, k = n / 2 + ind;
  const double f = (double)n;
  double x = 0.0;
  for (int i = 1; i <= k; ++i) {
    const int m = n + 1 - i;
    if (i == 1) x = 1.0 - 2.0 / ((f + 1.0) * f);
    if (i == 2) x = (z[n - 1] - 1.0) * 4.0 + z[n - 1];
    if (i == 3) x = (z[n - 2] - z[n - 1]) * 1.6 + z[n - 2];
    if (i > 3) x = (z[m] - z[m + 1]) * 3.0 + z[m + 2];
    if (i == k && ind == 1) x = 0.0;
    int niter = 0;
    double check = 1e-16, pa = 0.0, pb = 0.0, pc = 0.0;
    for (;;) {
      pb = 1.0;
      if (++niter > 100) check *= 10.0;
      pc = x;
      double dj = 1.0;
      for (int j = 2; j <= n; ++j) {
        dj += 1.0;
        pa = pb;
        pb = pc;
        pc = x * pb + (x * pb - pa) * (dj - 1.0) / dj;
      }
      pa = 1.0 / ((pb - x * pc) * f);
      pb = pa * pc * (1.0 - x * x);
      x = x - pb;
      if (!(fabs(pb) > check * fabs(x))) break;
    }
    z[m - 1] = x;
    w[m - 1] = 2.0 * pa * pa * (1.0 - x * x);
    if (i == k && ind == 1) continue;
    z[i - 1] = -z[m - 1];
    w[i - 1] = w[m - 1];
  }
}

// P^00, P^22, P^2-2, P^02 for l = 0..L-1 at u (GENER), into p[4][L].
static void gener(double u, int L, double* p1, double* p2, double* p3, double* p4) {
  for (int l = 0; l < L; ++l) p1[l] = p2[l] = p3[l] = p4[l] = 0.0;
  const double dup = 1.0 + u, dum = 1.0 - u, du = u * u, d6 = 0.25 * sqrt(6.0);
  p1[0] = 1.0;
  if (L > 1) p1[1] = u;
  if (L > 2) {
    p1[2] = 0.5 * (3.0 * du - 1.0);
    p2[2] = 0.25 * dup * dup;
    p3[2] = 0.25 * dum * dum;
    p4[2] = d6 * (du - 1.0);
  }
  for (int L1 = 3; L1 < L; ++L1) {
    const int l = L1 - 1;
    const double c1 = 1.0 / (l + 1), c2 = (double)(2 * l + 1);
    const double c3 = 1.0 / sqrt((double)((l + 1) * (l + 1) - 4)), c4 = sqrt((double)(l * l - 4));
    const double c5 = 1.0 / ((double)l * (double)((l + 1) * (l + 1) - 4));
    const double c6 = (double)(2 * l + 1) * (double)(l * (l + 1));
    const double c7 = (double)((2 * l + 1) * 4), c8 = (double)(l + 1) * (double)(l * l - 4);
    const double cu1 = c2 * u, cu2 = c6 * u, dl = (double)l;
    p1[L1] = c1 * (cu1 * p1[L1 - 1] - dl * p1[L1 - 2]);
    p2[L1] = c5 * ((cu2 - c7) * p2[L1 - 1] - c8 * p2[L1 - 2]);
    p3[L1] = c5 * ((cu2 + c7) * p3[L1 - 1] - c8 * p3[L1 - 2]);
    p4[L1] = c3 * (cu1 * p4[L1 - 1] - c4 * p4[L1 - 2]);
  }
}

// Host construction of G[4][n_ang][ng] (layout t, angle, l).
static void build_tables(int ng, int n_ang, const double* ang_deg, std::vector<double>& G) {
  const double d2r = acos(-1.0) / 180.0;
  std::vector<double> xx(n_ang);
  for (int a = 0; a < n_ang; ++a) xx[a] = ang_deg[a] * d2r;
  std::vector<double> z, w;
  gauss_nodes(ng, z, w);
  G.assign((size_t)4 * n_ang * ng, 0.0);
  std::vector<double> p(4 * (size_t)ng);
  for (int i = 0; i < ng; ++i) {
    const double x = acos(z[i]);
    // LINTERPOL (IEX = 1): interval (lo, hi) and weight t of hi
    int lo, hi;
    if (x < xx[0]) {
      lo = 0;
      hi = 1;
    } else if (x > xx[n_ang - 1]) {
      lo = n_ang - 2;
      hi = n_ang - 1;
    } else {
      int I = 1;
      while (I < n_ang && !(xx[I] > x)) ++I;
      if (I >= n_ang) I = n_ang - 1;  // x == last grid angle: weight 1 on it
      lo = I - 1;
      hi = I;
    }
    const double t = (x - xx[lo]) / (xx[hi] - xx[lo]);
    gener(z[i], ng, &p[0], &p[ng], &p[2 * (size_t)ng], &p[3 * (size_t)ng]);
    for (int tt = 0; tt < 4; ++tt) {
      double* Gt = &G[(size_t)tt * n_ang * ng];
      const double* pt = &p[(size_t)tt * ng];
      for (int l = 0; l < ng; ++l) {
        const double v = w[i] * pt[l];
        Gt[(size_t)lo * ng + l] += v * (1.0 - t);
        Gt[(size_t)hi * ng + l] += v * t;
      }
    }
  }
}

// The six angle sums of every entry as products on the FP64 matrix cores:
//   [entries x angles] x G_t[angles x l]
// for the rows F11, F44 (t = P00), F22 + F33 (P22), F22 - F33 (P2-2) and F12,
// F34 (P02).  F[e][6][n_ang] in table order p11, p12, p33, p34, p22, p44.  A
// wavefront takes 16 entries x 16 coefficients over one chunk of angles, four
// angles per v_mfma_f64_16x16x4_f64 (A[i][k] from lane i + 16k, B[k][j] from
// lane j + 16k, D[g + 4r][j] in lane j + 16g element r); the chunks split the
// angle sum so that a few entries still fill the GPU, and k_gsf_reduce adds
// them up in chunk order (deterministic).
//   part[chunk][e][c][l], c: 0 F11 P00, 1 (F22+F33) P22, 2 (F22-F33) P2-2,
//   3 F44 P00, 4 F12 P02, 5 F34 P02
typedef double double4_ __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(64) void k_gsf_expand(int64_t n_ent, int n_ang, int ng,
                                                   int chunk_ang, const double* __restrict__ G,
                                                   const double* __restrict__ F,
                                                   double* __restrict__ part) {
  const int lane = threadIdx.x, i = lane & 15, kq = lane >> 4;
  const int64_t e0 = (int64_t)blockIdx.x * 16;
  const int l0 = blockIdx.y * 16;
  const int ch = blockIdx.z;
  const int a_lo = ch * chunk_ang, a_hi = min(n_ang, a_lo + chunk_ang);
  const size_t plane = (size_t)n_ang * ng;
  const int64_t ea = e0 + i;
  const bool e_ok = ea < n_ent;
  const double* Fe = F + (size_t)(e_ok ? ea : 0) * 6 * n_ang;
  const int lb = l0 + i;
  const bool l_ok = lb < ng;
  double4_ acc[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) acc[c] = (double4_){0.0, 0.0, 0.0, 0.0};
  for (int a0 = a_lo; a0 < a_hi; a0 += 4) {
    const int a = a0 + kq;
    const bool ok = a < a_hi;
    double f11 = 0.0, f12 = 0.0, f33 = 0.0, f34 = 0.0, f22 = 0.0, f44 = 0.0;
    if (ok && e_ok) {
      f11 = Fe[a];
      f12 = Fe[n_ang + a];
      f33 = Fe[2 * n_ang + a];
      f34 = Fe[3 * n_ang + a];
      f22 = Fe[4 * n_ang + a];
      f44 = Fe[5 * n_ang + a];
    }
    double g1 = 0.0, g2 = 0.0, g3 = 0.0, g4 = 0.0;
    if (ok && l_ok) {
      const size_t o = (size_t)a * ng + lb;
      g1 = G[o];
      g2 = G[plane + o];
      g3 = G[2 * plane + o];
      g4 = G[3 * plane + o];
    }
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(f11, g1, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(f22 + f33, g2, acc[1], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(f22 - f33, g3, acc[2], 0, 0, 0);
    acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(f44, g1, acc[3], 0, 0, 0);
    acc[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(f12, g4, acc[4], 0, 0, 0);
    acc[5] = __builtin_amdgcn_mfma_f64_16x16x4f64(f34, g4, acc[5], 0, 0, 0);
  }
  // lane (i, kq) element r: entry e0 + kq + 4 r, coefficient l0 + i
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t e = e0 + kq + 4 * r;
      if (e < n_ent && l_ok) part[(((size_t)ch * n_ent + e) * 6 + c) * ng + lb] = acc[c][r];
    }
}

// raw[e][6][l] from the chunk partials, added in chunk order, then alpha /
// beta by the (l + 1/2) factors (spher_expan.f:335-347), in the table order
// p11 <- alpha1, p12 <- beta1, p33 <- alpha3, p34 <- beta2, p22 <- alpha2,
// p44 <- alpha4 (pmom(bin, wavelength, rh, p, m), convertncdf.py:76, 381-394).
__global__ __launch_bounds__(256) void k_gsf_reduce(int64_t n_ent, int ng, int n_chunks,
                                                    const double* __restrict__ part,
                                                    double* __restrict__ raw) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_ent * ng) return;
  const int64_t e = idx / ng;
  const int l = (int)(idx - e * ng);
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int ch = 0; ch < n_chunks; ++ch) {
    const double* pc = part + (((size_t)ch * n_ent + e) * 6) * ng + l;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] += pc[(size_t)c * ng];
  }
  const double cl = (double)l + 0.5;
  double* o = raw + (size_t)e * 6 * ng;
  const double a2 = acc[1] * cl * 0.5, a3 = acc[2] * cl * 0.5;
  o[0 * ng + l] = acc[0] * cl;   // p11 <- alpha1
  o[1 * ng + l] = acc[4] * cl;   // p12 <- beta1
  o[2 * ng + l] = a2 - a3;       // p33 <- alpha3
  o[3 * ng + l] = acc[5] * cl;   // p34 <- beta2
  o[4 * ng + l] = a2 + a3;       // p22 <- alpha2
  o[5 * ng + l] = acc[3] * cl;   // p44 <- alpha4
}

// Normalise by 1/alpha1(0) (spher_expan.f:95-103); optional 10-decimal rounding
// (the F17.10 output, :104-105).  One block per entry.
__global__ __launch_bounds__(256) void k_gsf_normalize(int ng, const double* __restrict__ raw,
                                                       double* __restrict__ out, int quantize) {
  const int64_t e = blockIdx.x;
  const double* r = raw + (size_t)e * 6 * ng;
  double* o = out + (size_t)e * 6 * ng;
  const double cnorm = (quantize & 2) ? 1.0 : 1.0 / r[0];
  for (int i = threadIdx.x; i < 6 * ng; i += blockDim.x) {
    double v = (quantize & 2) ? r[i] : r[i] * cnorm;
    if (quantize & 1) v = rint(v * 1e10) * 1e-10;
    o[i] = v;
  }
}

// MATR (spher_expan.f:419-517): the phase matrix at angles theta from its
// expansion coefficients, one thread per (entry, angle), by the program's
// upward recurrences of P^0_0, P^2_2, P^2_-2 and P^0_2 over l.
//   coef [e][6][l1max] in gm_gsf_expand's order (alpha1, beta1, alpha3, beta2,
//   alpha2, alpha4); out [e][6][n_ang] in the table order p11, p12, p33, p34,
//   p22, p44.
__global__ __launch_bounds__(64) void k_gsf_matr(int l1max, const double* __restrict__ coef,
                                                 int n_ang, const double* __restrict__ theta,
                                                 double* __restrict__ out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = blockIdx.y;
  if (a >= n_ang) return;
  const double* c = coef + (size_t)e * 6 * l1max;
  const double* A1 = c;
  const double* B1 = c + l1max;
  const double* A3 = c + 2 * l1max;
  const double* B2 = c + 3 * l1max;
  const double* A2 = c + 4 * l1max;
  const double* A4 = c + 5 * l1max;
  const int lmax = l1max - 1;
  const double d6 = sqrt(6.0) * 0.25;
  const double u = cos(theta[a]);
  double f11 = 0.0, f2 = 0.0, f3 = 0.0, f44 = 0.0, f12 = 0.0, f34 = 0.0;
  double p1 = 0.0, p2 = 0.0, p3 = 0.0, p4 = 0.0;
  double pp1 = 1.0;
  double pp2 = 0.25 * (1.0 + u) * (1.0 + u);
  double pp3 = 0.25 * (1.0 - u) * (1.0 - u);
  double pp4 = d6 * (u * u - 1.0);
  for (int l1 = 1; l1 <= l1max; ++l1) {
    const int l = l1 - 1;
    const double dl = (double)l, dl1 = (double)l1;
    f11 = f11 + A1[l] * pp1;
    f44 = f44 + A4[l] * pp1;
    const double pl1 = (double)(2 * l + 1);
    if (l != lmax) {
      const double pn = (pl1 * u * pp1 - dl * p1) / dl1;
      p1 = pp1;
      pp1 = pn;
    }
    if (l < 2) continue;
    f2 = f2 + (A2[l] + A3[l]) * pp2;
    f3 = f3 + (A2[l] - A3[l]) * pp3;
    f12 = f12 + B1[l] * pp4;
    f34 = f34 + B2[l] * pp4;
    if (l == lmax) continue;
    const double pl2 = dl * dl1 * u;
    const double pl3 = dl1 * (dl * dl - 4.0);
    const double pl4 = 1.0 / (dl * (dl1 * dl1 - 4.0));
    double pn = (pl1 * (pl2 - 4.0) * pp2 - pl3 * p2) * pl4;
    p2 = pp2;
    pp2 = pn;
    pn = (pl1 * (pl2 + 4.0) * pp3 - pl3 * p3) * pl4;
    p3 = pp3;
    pp3 = pn;
    pn = (pl1 * u * pp4 - sqrt(dl * dl - 4.0) * p4) / sqrt(dl1 * dl1 - 4.0);
    p4 = pp4;
    pp4 = pn;
  }
  double* o = out + (size_t)e * 6 * n_ang;
  o[a] = f11;
  o[n_ang + a] = f12;
  o[2 * n_ang + a] = (f2 - f3) * 0.5;   // F33
  o[3 * n_ang + a] = f34;
  o[4 * n_ang + a] = (f2 + f3) * 0.5;   // F22
  o[5 * n_ang + a] = f44;
}

// Legendre series sum_l c[l] P_l(x) (convertncdf.py:32-43), one thread per x.
__global__ void k_legendre_eval(int64_t n_x, const double* __restrict__ x, int n_series, int L,
                                const double* __restrict__ coef, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int s = blockIdx.y;
  if (i >= n_x || s >= n_series) return;
  const double u = x[i];
  const double* c = coef + (size_t)s * L;
  double pm = 1.0, p = u;
  double acc = c[0];
  if (L > 1) acc += c[1] * u;
  for (int m = 2; m < L; ++m) {
    double pn = 1.0 / m * ((2.0 * m - 1.0) * u * p - (m - 1.0) * pm);
    acc += c[m] * pn;
    pm = p;
    p = pn;
  }
  out[(size_t)s * n_x + i] = acc;
}

// Make ctx->gsf_tab hold G[4][n_ang][ng] for this (ng, angle grid); rebuilt
// only when either changes.
int gsf_tables(gm_ctx* ctx, int ng, int n_ang, const double* ang_deg_host, hipStream_t s) {
  double key = 0.0;
  for (int a = 0; a < n_ang; ++a) key += ang_deg_host[a] * (1.0 + 1e-3 * a);
  if (ctx->gsf_ng == ng && ctx->gsf_nang == n_ang && ctx->gsf_key == key) return GM_OK;
  std::vector<double> G;
  build_tables(ng, n_ang, ang_deg_host, G);
  int rc = ctx->reserve(ctx->gsf_tab, G.size() * sizeof(double));
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(ctx->gsf_tab.ptr, G.data(), G.size() * sizeof(double),
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // G is a host temporary
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "gsf tables: %s", hipGetErrorString(e));
  ctx->gsf_ng = ng;
  ctx->gsf_nang = n_ang;
  ctx->gsf_key = key;
  return GM_OK;
}

}  // namespace gm

extern "C" {

int gm_gsf_expand(gm_ctx* ctx, int64_t n_entries, int n_ang, const double* ang_deg_host,
                  const double* F, int ng, int quantize, double* out, gm_stream stream) {
  if (!ctx || n_entries < 0 || n_ang < 2 || ng < 3 || !ang_deg_host)
    return gm_fail(-1, "gm_gsf_expand: bad arguments");
  if (n_entries == 0) return 0;
  if (!F || !out) return gm_fail(-1, "gm_gsf_expand: NULL array");
  for (int a = 1; a < n_ang; ++a)
    if (!(ang_deg_host[a] > ang_deg_host[a - 1]))
      return gm_fail(-1, "gm_gsf_expand: angles must increase");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  int rc = gm::gsf_tables(ctx, ng, n_ang, ang_deg_host, s);
  if (rc) return rc;
  // angle chunks (>= 32 angles each) until ~2048 wavefronts are in flight
  const int64_t tiles = ((n_entries + 15) / 16) * ((ng + 15) / 16);
  int n_chunks = (int)std::max<int64_t>(
      1, std::min<int64_t>((2048 + tiles - 1) / tiles, (n_ang + 31) / 32));
  const int chunk_ang = ((n_ang + n_chunks - 1) / n_chunks + 3) / 4 * 4;
  n_chunks = (n_ang + chunk_ang - 1) / chunk_ang;
  if ((n_entries + 15) / 16 > 0x7fffffffLL || (ng + 15) / 16 > 65535)
    return gm_fail(-1, "gm_gsf_expand: too many entries or coefficients for one call");
  const size_t raw_bytes = sizeof(double) * (size_t)n_entries * 6 * ng;
  rc = ctx->reserve(ctx->gemm_tmp, raw_bytes * (1 + (size_t)n_chunks));
  if (rc) return rc;
  double* raw = (double*)ctx->gemm_tmp.ptr;
  double* part = raw + (size_t)n_entries * 6 * ng;
  size_t slot;
  ctx->mark_begin(gm::kKGsf, s, &slot);
  hipLaunchKernelGGL(gm::k_gsf_expand,
                     dim3((unsigned)((n_entries + 15) / 16), (unsigned)((ng + 15) / 16),
                          (unsigned)n_chunks),
                     dim3(64), 0, s, n_entries, n_ang, ng, chunk_ang,
                     (const double*)ctx->gsf_tab.ptr, F, part);
  hipLaunchKernelGGL(gm::k_gsf_reduce, dim3((unsigned)((n_entries * ng + 255) / 256)), dim3(256),
                     0, s, n_entries, ng, n_chunks, (const double*)part, raw);
  hipLaunchKernelGGL(gm::k_gsf_normalize, dim3((unsigned)n_entries), dim3(256), 0, s, ng, raw, out,
                     quantize);
  ctx->mark_end(slot, s);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "k_gsf_expand: %s", hipGetErrorString(e));
  return 0;
}

int gm_gsf_matr(gm_ctx* ctx, int64_t n_entries, int l1max, const double* coef, int n_ang,
                const double* theta_rad, double* out, gm_stream stream) {
  if (!ctx || n_entries < 0 || n_entries > 65535 || l1max < 3 || n_ang < 1 || !coef ||
      !theta_rad || !out)
    return gm_fail(-1, "gm_gsf_matr: bad arguments");
  if (n_entries == 0) return 0;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(gm::k_gsf_matr, dim3((unsigned)((n_ang + 63) / 64), (unsigned)n_entries),
                     dim3(64), 0, (hipStream_t)stream, l1max, coef, n_ang, theta_rad, out);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "gm_gsf_matr: %s", hipGetErrorString(e));
  return 0;
}

int gm_legendre_eval(gm_ctx* ctx, int64_t n_x, const double* x, int n_series, int L,
                     const double* coef, double* out, gm_stream stream) {
  if (!ctx || n_x < 0 || n_series < 0 || L < 1) return gm_fail(-1, "gm_legendre_eval: bad arguments");
  if (n_x == 0 || n_series == 0) return 0;
  if (!x || !coef || !out) return gm_fail(-1, "gm_legendre_eval: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(1, "hipSetDevice: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(gm::k_legendre_eval, dim3((unsigned)((n_x + 127) / 128), n_series), dim3(128),
                     0, (hipStream_t)stream, n_x, x, n_series, L, coef, out);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "k_legendre_eval: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
