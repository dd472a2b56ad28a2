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
//
// Reduced-precision paths (the coefficients and the Mie series sums stay FP64):
//   FP32    v_mfma_f32_16x16x4_f32 on FP32 copies of U, V, P, M.
//   BF16X3  v_mfma_f32_16x16x32_bf16 on split operands x = x_hi + x_lo (both
//           BF16, x_lo = bf16(x - x_hi)); U P ~ U_hi P_hi + U_hi P_lo + U_lo P_hi
//           (the lo*lo term, ~2^-18 relative, is dropped), FP32 accumulation.
// Both read P, M from a per-call table (k_s12_prep_tables) padded to whole
// K-steps and 64-angle tiles, so the inner loops carry no bounds tests on it.
// f32 / bf16 16x16 results land as D[4g + r][j] in lane j + 16g, element r;
// A row i is then point p0 + (i >> 1), component i & 1, which gives the same
// per-lane element order as the FP64 kernel: (0,1) = point p0 + 2g (re, im),
// (2,3) = point p0 + 2g + 1.
#include "gm_complex.h"
#include "gm_ctx.h"

typedef double double4_ __attribute__((ext_vector_type(4)));
typedef float float4_ __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_ __attribute__((ext_vector_type(8)));

namespace gm {

enum { kOutS12 = 0, kOutMueller = 1, kOutP11 = 2 };

// S1 = (sp + sm) / 2, S2 = (sp - sm) / 2 for point p, angle j; then either the
// amplitudes or the Mueller elements (dointegration.py:1045-1050).
__device__ __forceinline__ void s12_store(int mode, int64_t p, int j, int n_ang, cd sp, cd sm,
                                          double* __restrict__ out0, double* __restrict__ out1) {
  cd s1 = (sp + sm) * 0.5;
  cd s2 = (sp - sm) * 0.5;
  size_t o = (size_t)p * n_ang + j;
  if (mode == kOutS12) {
    out0[2 * o] = s1.re;
    out0[2 * o + 1] = s1.im;
    out1[2 * o] = s2.re;
    out1[2 * o + 1] = s2.im;
  } else if (mode == kOutP11) {
    out0[o] = 0.5 * (norm2(s1) + norm2(s2));   // s11 only, layout [point][angle]
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

// MT = M-tiles (8 points each) per wavefront: with MT = 2 every angle-table
// load feeds twice as many MFMAs.
template <int MT>
__global__ __launch_bounds__(256) void k_s12_f64(
    int64_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ nmax,
    const double* __restrict__ an, const double* __restrict__ bn, int n_ang,
    const double* __restrict__ pi_t, const double* __restrict__ tau_t, int nmax_cap, int mode,
    int wave_points, double* __restrict__ out0, double* __restrict__ out1) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  // 64 angles x 8 MT points per wavefront.  The waves of a block take
  // different angles of the same points, or (wave_points) different points of
  // the same angles, whose angle-table rows they then read through one CU's
  // cache.
  const int64_t p0 = wave_points ? ((int64_t)blockIdx.x * nw + wave) * 8 * MT
                                 : (int64_t)blockIdx.x * 8 * MT;
  const int j0 = wave_points ? blockIdx.y * 64 : (blockIdx.y * nw + wave) * 64;
  if (j0 >= n_ang) return;  // whole wavefront past the last angle
  const int i = lane & 15;   // A row / B column / D column
  const int kq = lane >> 4;  // k within the 4-step, D row group
  // The f64 16x16x4 result row of (lane group g, element r) is g + 4r.  A row
  // i = g + 4r of M-tile mt is assigned to point p0 + 8 mt + 2g + (r >> 1),
  // component r & 1, so lane group g ends up holding (re, im) of points
  // p0 + 8 mt + 2g and p0 + 8 mt + 2g + 1.
  const int comp = (i >> 2) & 1;
  int nm_a[MT];
  int64_t oa[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int64_t pa = p0 + 8 * mt + 2 * (i & 3) + (i >> 3);
    const bool ok = pa < n;
    nm_a[mt] = ok ? nmax[pa] : 0;
    oa[mt] = ok ? off[pa] : 0;
  }
  // tile K = max nmax of the 8 MT points
  int kmax = 0;
  for (int q = 0; q < 8 * MT; ++q) {
    int64_t pq = p0 + q;
    if (pq < n) kmax = max(kmax, nmax[pq]);
  }
  kmax = min(kmax, nmax_cap);
  double4_ accp[MT][4], accm[MT][4];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      accp[mt][t] = (double4_){0.0, 0.0, 0.0, 0.0};
      accm[mt][t] = (double4_){0.0, 0.0, 0.0, 0.0};
    }
  // Operands of step k0 + 4 are loaded while the MFMAs of step k0 run
  // (register double buffer), so the loads' latency overlaps the products.
  double ar_n[MT], br_n[MT], pp_n[4], tt_n[4];
  auto load_step = [&](int k0) {
    const int kk = k0 + kq;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      ar_n[mt] = 0.0;
      br_n[mt] = 0.0;
      if (kk < nm_a[mt]) {
        ar_n[mt] = an[2 * (oa[mt] + kk) + comp];
        br_n[mt] = bn[2 * (oa[mt] + kk) + comp];
      }
    }
    const bool krow = kk < kmax;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int j = j0 + t * 16 + i;
      pp_n[t] = 0.0;
      tt_n[t] = 0.0;
      if (krow && j < n_ang) {
        pp_n[t] = pi_t[(size_t)kk * n_ang + j];
        tt_n[t] = tau_t[(size_t)kk * n_ang + j];
      }
    }
  };
  if (kmax > 0) load_step(0);
  for (int k0 = 0; k0 < kmax; k0 += 4) {
    double ua[MT], va[MT], P[4], M[4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      ua[mt] = ar_n[mt] + br_n[mt];
      va[mt] = ar_n[mt] - br_n[mt];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      P[t] = pp_n[t] + tt_n[t];
      M[t] = pp_n[t] - tt_n[t];
    }
    if (k0 + 4 < kmax) load_step(k0 + 4);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        accp[mt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[mt], P[t], accp[mt][t], 0, 0, 0);
        accm[mt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[mt], M[t], accm[mt][t], 0, 0, 0);
      }
    }
  }
  // epilogue: elements (0,1) = point p0+8mt+2kq (re,im), (2,3) = the next point
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int j = j0 + t * 16 + i;
      if (j >= n_ang) continue;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t p = p0 + 8 * mt + 2 * kq + h;
        if (p >= n) continue;
        s12_store(mode, p, j, n_ang, cd{accp[mt][t][2 * h], accp[mt][t][2 * h + 1]},
                  cd{accm[mt][t][2 * h], accm[mt][t][2 * h + 1]}, out0, out1);
      }
    }
}

// LDS-staged variant: the kS12LdsWaves wavefronts of a block take 8 points each
// at the same 64 angles and share the angle tables: every stage of
// kS12LdsRows K-rows of P = pi' + tau' and M = pi' - tau' (64 angles) is read
// from global memory once per block into LDS, while the previous stage's
// MFMAs run (the next stage is held in registers, then stored behind a
// barrier).  The K loop runs to the block's largest n_max; rows past a
// point's own n_max multiply zero coefficients.
constexpr int kS12LdsWaves = 4;
constexpr int kS12LdsRows = 16;
constexpr int kS12LdsStride = 64 + 2;   // doubles per LDS row (padding against bank conflicts)
template <int MT>
__global__ __launch_bounds__(64 * kS12LdsWaves) void k_s12_f64_lds(
    int64_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ nmax,
    const double* __restrict__ an, const double* __restrict__ bn, int n_ang,
    const double* __restrict__ pi_t, const double* __restrict__ tau_t, int nmax_cap, int mode,
    double* __restrict__ out0, double* __restrict__ out1) {
  __shared__ double Ps[2][kS12LdsRows][kS12LdsStride];
  __shared__ double Ms[2][kS12LdsRows][kS12LdsStride];
  __shared__ int kmax_w[kS12LdsWaves];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t p0 = ((int64_t)blockIdx.x * kS12LdsWaves + wave) * 8 * MT;
  const int j0 = blockIdx.y * 64;
  const int i = lane & 15;
  const int kq = lane >> 4;
  const int comp = (i >> 2) & 1;
  int nm_a[MT];
  int64_t oa[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int64_t pa = p0 + 8 * mt + 2 * (i & 3) + (i >> 3);
    const bool ok = pa < n;
    nm_a[mt] = ok ? nmax[pa] : 0;
    oa[mt] = ok ? off[pa] : 0;
  }
  int kw = 0;
  for (int q = 0; q < 8 * MT; ++q) {
    const int64_t pq = p0 + q;
    if (pq < n) kw = max(kw, nmax[pq]);
  }
  if (lane == 0) kmax_w[wave] = min(kw, nmax_cap);
  __syncthreads();
  int kmax = 0;
  for (int w = 0; w < kS12LdsWaves; ++w) kmax = max(kmax, kmax_w[w]);
  // cooperative stage loads: thread tid covers (row, column) pairs tid + 256 e
  constexpr int kPer = kS12LdsRows * 64 / (64 * kS12LdsWaves);
  double pn[kPer], mn[kPer];
  auto load_stage = [&](int s0) {
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const int q = tid + 64 * kS12LdsWaves * e;
      const int r = q >> 6, c = q & 63;
      const int kk = s0 + r, j = j0 + c;
      double pp = 0.0, tt = 0.0;
      if (kk < kmax && j < n_ang) {
        pp = pi_t[(size_t)kk * n_ang + j];
        tt = tau_t[(size_t)kk * n_ang + j];
      }
      pn[e] = pp + tt;
      mn[e] = pp - tt;
    }
  };
  auto store_stage = [&](int buf) {
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const int q = tid + 64 * kS12LdsWaves * e;
      const int r = q >> 6, c = q & 63;
      Ps[buf][r][c] = pn[e];
      Ms[buf][r][c] = mn[e];
    }
  };
  double4_ accp[MT][4], accm[MT][4];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      accp[mt][t] = (double4_){0.0, 0.0, 0.0, 0.0};
      accm[mt][t] = (double4_){0.0, 0.0, 0.0, 0.0};
    }
  if (kmax > 0) {
    load_stage(0);
    store_stage(0);
  }
  __syncthreads();
  int buf = 0;
  for (int s0 = 0; s0 < kmax; s0 += kS12LdsRows) {
    const bool more = s0 + kS12LdsRows < kmax;
    if (more) load_stage(s0 + kS12LdsRows);     // in flight during this stage's MFMAs
#pragma unroll
    for (int ks = 0; ks < kS12LdsRows / 4; ++ks) {
      const int kk = s0 + 4 * ks + kq;
      double ua[MT], va[MT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        ua[mt] = 0.0;
        va[mt] = 0.0;
        if (kk < nm_a[mt]) {
          const double ar = an[2 * (oa[mt] + kk) + comp];
          const double br = bn[2 * (oa[mt] + kk) + comp];
          ua[mt] = ar + br;
          va[mt] = ar - br;
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double P = Ps[buf][4 * ks + kq][t * 16 + i];
        const double M = Ms[buf][4 * ks + kq][t * 16 + i];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          accp[mt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[mt], P, accp[mt][t], 0, 0, 0);
          accm[mt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[mt], M, accm[mt][t], 0, 0, 0);
        }
      }
    }
    if (more) store_stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int j = j0 + t * 16 + i;
      if (j >= n_ang) continue;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t p = p0 + 8 * mt + 2 * kq + h;
        if (p >= n) continue;
        s12_store(mode, p, j, n_ang, cd{accp[mt][t][2 * h], accp[mt][t][2 * h + 1]},
                  cd{accm[mt][t][2 * h], accm[mt][t][2 * h + 1]}, out0, out1);
      }
    }
}

// ---- S1/S2 -> Mueller -> weighted point sums, fused --------------------------
//
// integratePSD integrates each Mueller element of rawMie against a row of PSD
// weights (dointegration.py:1044-1050, 1164-1166: np.dot(s_ij.T, psd)).  When
// every weight row covers a contiguous range of points and few rows share a
// point (the runoptics passes: one block of points per (rh, refractive index),
// one row per mode), the Mueller planes need not exist in memory: each
// wavefront forms S1/S2 of 8-point tiles x 64 angles on the matrix cores as
// k_s12_f64 does, turns them into s11, s12, s33, s34 in registers and adds
// weight x element into per-row sums.  A block takes one chunk (one or two
// whole 8-point tiles) of one row group (rows with the same range), its four
// wavefronts taking 64 angles each; the sums go to a partial buffer
// [chunk][RM][4][n_ang] that k_s12_rows_reduce adds up per row in chunk
// order (deterministic).  Memory traffic drops from 2 x 32 B per (point,
// angle) to the partials (8 / RM times less per tile).  Chunks of one or two
// tiles keep the work units as small as k_s12_f64's, so the long K loops of
// large size parameters do not end up in a few long waves.
constexpr int kRowsPerCall = 64;   // rows per call (range arrays passed by value)
constexpr int kRowsWaves = 4;      // wavefronts (64 angles each) per block

struct RowGroups {
  int n_groups;
  int64_t start[kRowsPerCall], count[kRowsPerCall];
  int row0[kRowsPerCall], nrow[kRowsPerCall];
  int64_t chunk0[kRowsPerCall + 1];   // first chunk of each group; chunk0[n_groups] = total
};

// RM = most rows of one group this instance accumulates (registers: 16 RM
// doubles per lane for the sums).  MT = 8-point tiles multiplied together:
// each angle-table word a lane loads feeds the products of all MT tiles (the
// table loads, not the matrix cores, bound this kernel at MT = 1); a chunk
// of chunk_pts points runs as consecutive groups of MT tiles.
template <int RM, int MT>
__global__ __launch_bounds__(64 * kRowsWaves) void k_s12_rows_f64(
    RowGroups g, int chunk_pts, const int64_t* __restrict__ off, const int32_t* __restrict__ nmax,
    const double* __restrict__ an, const double* __restrict__ bn, int n_ang,
    const double* __restrict__ pi_t, const double* __restrict__ tau_t, int nmax_cap,
    const double* __restrict__ w, int64_t ldw, double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t chunk = blockIdx.x;
  const int j0 = (blockIdx.y * kRowsWaves + (threadIdx.x >> 6)) * 64;
  if (j0 >= n_ang) return;   // whole wavefront past the last angle (no barriers below)
  int gi = 0;
  while (gi + 1 < g.n_groups && g.chunk0[gi + 1] <= chunk) ++gi;
  const int64_t cstart = g.start[gi] + (chunk - g.chunk0[gi]) * chunk_pts;
  const int64_t cend = min(cstart + (int64_t)chunk_pts, g.start[gi] + g.count[gi]);
  const int nr = g.nrow[gi];
  const double* wr = w + (size_t)g.row0[gi] * ldw;
  const int i = lane & 15;
  const int kq = lane >> 4;
  const int comp = (i >> 2) & 1;
  double sum[RM][4][4];   // [row][angle sub-tile t][element]
#pragma unroll
  for (int r = 0; r < RM; ++r)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) sum[r][t][e] = 0.0;
  for (int64_t p0 = cstart; p0 < cend; p0 += 8 * MT) {
    int nm_a[MT];
    int64_t oa[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int64_t pa = p0 + 8 * mt + 2 * (i & 3) + (i >> 3);
      const bool ok = pa < cend;
      nm_a[mt] = ok ? nmax[pa] : 0;
      oa[mt] = ok ? off[pa] : 0;
    }
    int kmax = 0;
    for (int q = 0; q < 8 * MT; ++q)
      if (p0 + q < cend) kmax = max(kmax, nmax[p0 + q]);
    kmax = min(kmax, nmax_cap);
    double4_ accp[MT][4], accm[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        accp[mt][t] = (double4_){0.0, 0.0, 0.0, 0.0};
        accm[mt][t] = (double4_){0.0, 0.0, 0.0, 0.0};
      }
    double ar_n[MT], br_n[MT], pp_n[4], tt_n[4];
    auto load_step = [&](int k0) {
      const int kk = k0 + kq;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        ar_n[mt] = 0.0;
        br_n[mt] = 0.0;
        if (kk < nm_a[mt]) {
          ar_n[mt] = an[2 * (oa[mt] + kk) + comp];
          br_n[mt] = bn[2 * (oa[mt] + kk) + comp];
        }
      }
      const bool krow = kk < kmax;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int j = j0 + t * 16 + i;
        pp_n[t] = 0.0;
        tt_n[t] = 0.0;
        if (krow && j < n_ang) {
          pp_n[t] = pi_t[(size_t)kk * n_ang + j];
          tt_n[t] = tau_t[(size_t)kk * n_ang + j];
        }
      }
    };
    if (kmax > 0) load_step(0);
    for (int k0 = 0; k0 < kmax; k0 += 4) {
      double ua[MT], va[MT], P[4], M[4];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        ua[mt] = ar_n[mt] + br_n[mt];
        va[mt] = ar_n[mt] - br_n[mt];
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        P[t] = pp_n[t] + tt_n[t];
        M[t] = pp_n[t] - tt_n[t];
      }
      if (k0 + 4 < kmax) load_step(k0 + 4);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          accp[mt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[mt], P[t], accp[mt][t], 0, 0, 0);
          accm[mt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[mt], M[t], accm[mt][t], 0, 0, 0);
        }
    }
    // lane (i, kq) holds angle j0 + 16 t + i of points p0 + 8 mt + 2 kq + h
    // (h = 0, 1)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t p = p0 + 8 * mt + 2 * kq + h;
        double wv[RM];
#pragma unroll
        for (int r = 0; r < RM; ++r)
          wv[r] = (r < nr && p < cend) ? wr[(size_t)r * ldw + p] : 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const cd sp = cd{accp[mt][t][2 * h], accp[mt][t][2 * h + 1]};
          const cd sm = cd{accm[mt][t][2 * h], accm[mt][t][2 * h + 1]};
          const cd s1 = (sp + sm) * 0.5, s2 = (sp - sm) * 0.5;
          const double a1 = norm2(s1), a2 = norm2(s2);
          const cd c = s1 * conj(s2);
          const double el[4] = {0.5 * (a1 + a2), 0.5 * (a2 - a1), c.re, c.im};
#pragma unroll
          for (int r = 0; r < RM; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) sum[r][t][e] = fma(wv[r], el[e], sum[r][t][e]);
        }
      }
  }
  // add the four point pairs (lanes i + 16 kq) and write; lane group kq
  // writes element kq
  double* dst = part + (size_t)chunk * RM * 4 * n_ang;
#pragma unroll
  for (int r = 0; r < RM; ++r)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      double mine = 0.0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double v = sum[r][t][e];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (e == kq) mine = v;
      }
      const int j = j0 + t * 16 + i;
      if (r < nr && j < n_ang) dst[((size_t)r * 4 + kq) * n_ang + j] = mine;
    }
}

// out[row][e][j] = sum over the row group's chunks of the partials.  A block
// covers 64 consecutive (e, j) of one row; its 16 wavefronts take every 16th
// chunk each, and the 16 sums are added in wavefront order (deterministic).
constexpr int kReduceWaves = 16;
template <int RM>
__global__ __launch_bounds__(64 * kReduceWaves) void k_s12_rows_reduce(
    RowGroups g, int n_ang, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[kReduceWaves][64];
  const int lane = threadIdx.x & 63;
  const int q = threadIdx.x >> 6;
  const int per_row = 4 * n_ang;
  const int slabs = (per_row + 63) / 64;
  const int row = (int)(blockIdx.x / slabs);
  const int ej = (int)(blockIdx.x % slabs) * 64 + lane;
  int gi = 0;
  while (gi + 1 < g.n_groups && g.row0[gi + 1] <= row) ++gi;
  const int rl = row - g.row0[gi];
  double s = 0.0;
  if (ej < per_row)
    for (int64_t c = g.chunk0[gi] + q; c < g.chunk0[gi + 1]; c += kReduceWaves)
      s += part[((size_t)c * RM + rl) * per_row + ej];
  red[q][lane] = s;
  __syncthreads();
  if (q == 0 && ej < per_row) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kReduceWaves; ++k) t += red[k][lane];
    out[(size_t)row * per_row + ej] = t;
  }
}

// ---- reduced-precision paths -------------------------------------------------

constexpr int kLpK = 32;     // K padding of the reduced-precision tables
constexpr int kLpAng = 64;   // angle padding

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) {
  return __uint_as_float((unsigned)h << 16);
}
// x = hi + lo, both BF16 (hi = bf16(x), lo = bf16(x - hi) with the residual in FP64).
__device__ __forceinline__ void bf16_split(double x, unsigned short& hi, unsigned short& lo) {
  hi = bf16_rne((float)x);
  lo = bf16_rne((float)(x - (double)bf16_val(hi)));
}

// P = pi + tau and M = pi - tau from the FP64 tables [k][angle] (k < nmax_cap).
//   FP32 : t32[2][kpad][apad]            (plane 0 P, plane 1 M)
//   BF16 : t16[4][kpad / 8][apad][8]     (planes P_hi, P_lo, M_hi, M_lo)
// Zero outside k < nmax_cap, j < n_ang.
__global__ void k_s12_prep_tables(int n_ang, int nmax_cap, int kpad, int apad,
                                  const double* __restrict__ pi_t,
                                  const double* __restrict__ tau_t, int precision,
                                  float* __restrict__ t32, unsigned short* __restrict__ t16) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)kpad * apad) return;
  const int k = (int)(idx / apad), j = (int)(idx % apad);
  double P = 0.0, M = 0.0;
  if (k < nmax_cap && j < n_ang) {
    const double pv = pi_t[(size_t)k * n_ang + j], tv = tau_t[(size_t)k * n_ang + j];
    P = pv + tv;
    M = pv - tv;
  }
  if (precision == GM_PREC_FP32) {
    t32[(size_t)k * apad + j] = (float)P;
    t32[((size_t)kpad + k) * apad + j] = (float)M;
  } else {
    const size_t plane = (size_t)kpad * apad;
    const size_t o = ((size_t)(k >> 3) * apad + j) * 8 + (k & 7);
    unsigned short h, l;
    bf16_split(P, h, l);
    t16[o] = h;
    t16[plane + o] = l;
    bf16_split(M, h, l);
    t16[2 * plane + o] = h;
    t16[3 * plane + o] = l;
  }
}

// K extent of the tile of 8 points starting at p0, rounded up to a whole step.
__device__ __forceinline__ int tile_kmax(int64_t n, int64_t p0, const int32_t* __restrict__ nmax,
                                         int nmax_cap, int step) {
  int kmax = 0;
  for (int q = 0; q < 8; ++q) {
    int64_t pq = p0 + q;
    if (pq < n) kmax = max(kmax, nmax[pq]);
  }
  kmax = min(kmax, nmax_cap);
  return (kmax + step - 1) / step * step;
}

__global__ __launch_bounds__(256) void k_s12_f32(
    int64_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ nmax,
    const double* __restrict__ an, const double* __restrict__ bn, int n_ang, int apad, int kpad,
    const float* __restrict__ t32, int nmax_cap, int mode, double* __restrict__ out0,
    double* __restrict__ out1) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t p0 = (int64_t)blockIdx.x * 8;
  const int j0 = (blockIdx.y * (blockDim.x >> 6) + wave) * 64;  // 64 angles per wavefront
  if (j0 >= n_ang) return;
  const int i = lane & 15, kq = lane >> 4;
  const int64_t pa = p0 + (i >> 1);
  const int comp = i & 1;
  const bool pa_ok = pa < n;
  const int nm_a = pa_ok ? nmax[pa] : 0;
  const int64_t oa = pa_ok ? off[pa] : 0;
  const int kmax = tile_kmax(n, p0, nmax, nmax_cap, 4);
  const float* Pt = t32;
  const float* Mt = t32 + (size_t)kpad * apad;
  float4_ accp[4], accm[4];
  for (int t = 0; t < 4; ++t) {
    accp[t] = (float4_){0.f, 0.f, 0.f, 0.f};
    accm[t] = (float4_){0.f, 0.f, 0.f, 0.f};
  }
  // register double buffer: step k0 + 4's operands load during step k0's MFMAs
  double ar_n = 0.0, br_n = 0.0;
  float p_n[4], m_n[4];
  auto load_step = [&](int k0) {
    const int kk = k0 + kq;
    ar_n = 0.0;
    br_n = 0.0;
    if (kk < nm_a) {
      ar_n = an[2 * (oa + kk) + comp];
      br_n = bn[2 * (oa + kk) + comp];
    }
    const size_t row = (size_t)kk * apad + j0 + i;   // rows < kpad: zero padded
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      p_n[t] = Pt[row + 16 * t];
      m_n[t] = Mt[row + 16 * t];
    }
  };
  if (kmax > 0) load_step(0);
  for (int k0 = 0; k0 < kmax; k0 += 4) {
    const float ua = (float)(ar_n + br_n), va = (float)(ar_n - br_n);
    float P[4], M[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      P[t] = p_n[t];
      M[t] = m_n[t];
    }
    if (k0 + 4 < kmax) load_step(k0 + 4);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      accp[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ua, P[t], accp[t], 0, 0, 0);
      accm[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(va, M[t], accm[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int j = j0 + t * 16 + i;
    if (j >= n_ang) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t p = p0 + 2 * kq + h;
      if (p >= n) continue;
      s12_store(mode, p, j, n_ang, cd{(double)accp[t][2 * h], (double)accp[t][2 * h + 1]},
                cd{(double)accm[t][2 * h], (double)accm[t][2 * h + 1]}, out0, out1);
    }
  }
}

__global__ __launch_bounds__(256) void k_s12_bf16x3(
    int64_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ nmax,
    const double* __restrict__ an, const double* __restrict__ bn, int n_ang, int apad, int kpad,
    const unsigned short* __restrict__ t16, int nmax_cap, int mode, double* __restrict__ out0,
    double* __restrict__ out1) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t p0 = (int64_t)blockIdx.x * 8;
  const int j0 = (blockIdx.y * (blockDim.x >> 6) + wave) * 64;  // 64 angles per wavefront
  if (j0 >= n_ang) return;
  const int i = lane & 15, kq = lane >> 4;
  const int64_t pa = p0 + (i >> 1);
  const int comp = i & 1;
  const bool pa_ok = pa < n;
  const int nm_a = pa_ok ? nmax[pa] : 0;
  const int64_t oa = pa_ok ? off[pa] : 0;
  const int kmax = tile_kmax(n, p0, nmax, nmax_cap, kLpK);
  const size_t plane = (size_t)kpad * apad;
  const uint4* tab = (const uint4*)t16;   // 8 BF16 (one K-octet) per uint4
  const size_t plane4 = plane / 8;
  float4_ accp[4], accm[4];
  for (int t = 0; t < 4; ++t) {
    accp[t] = (float4_){0.f, 0.f, 0.f, 0.f};
    accm[t] = (float4_){0.f, 0.f, 0.f, 0.f};
  }
  // register double buffer of the next K-step's table words (the coefficient
  // loads are issued before the current step's products as well)
  uint4 tn[4][4];
  auto load_tab = [&](int k0) {
    const size_t o4 = (size_t)((k0 + 8 * kq) >> 3) * apad + j0 + i;   // rows < kpad: zero padded
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int pl4 = 0; pl4 < 4; ++pl4) tn[t][pl4] = tab[pl4 * plane4 + o4 + 16 * t];
    }
  };
  if (kmax > 0) load_tab(0);
  for (int k0 = 0; k0 < kmax; k0 += kLpK) {
    const int kb = k0 + 8 * kq;   // this lane's K-octet
    double ar8[8], br8[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int kk = kb + t;
      ar8[t] = 0.0;
      br8[t] = 0.0;
      if (kk < nm_a) {
        ar8[t] = an[2 * (oa + kk) + comp];
        br8[t] = bn[2 * (oa + kk) + comp];
      }
    }
    uint4 tc[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int pl4 = 0; pl4 < 4; ++pl4) tc[t][pl4] = tn[t][pl4];
    if (k0 + kLpK < kmax) load_tab(k0 + kLpK);
    union { bf16x8_ v; unsigned short s[8]; } uh, ul, vh, vl;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      bf16_split(ar8[t] + br8[t], uh.s[t], ul.s[t]);
      bf16_split(ar8[t] - br8[t], vh.s[t], vl.s[t]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      union { uint4 q; bf16x8_ v; } ph, pl, mh, ml;
      ph.q = tc[t][0];
      pl.q = tc[t][1];
      mh.q = tc[t][2];
      ml.q = tc[t][3];
      accp[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uh.v, ph.v, accp[t], 0, 0, 0);
      accp[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uh.v, pl.v, accp[t], 0, 0, 0);
      accp[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ul.v, ph.v, accp[t], 0, 0, 0);
      accm[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh.v, mh.v, accm[t], 0, 0, 0);
      accm[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh.v, ml.v, accm[t], 0, 0, 0);
      accm[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vl.v, mh.v, accm[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int j = j0 + t * 16 + i;
    if (j >= n_ang) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t p = p0 + 2 * kq + h;
      if (p >= n) continue;
      s12_store(mode, p, j, n_ang, cd{(double)accp[t][2 * h], (double)accp[t][2 * h + 1]},
                cd{(double)accm[t][2 * h], (double)accm[t][2 * h + 1]}, out0, out1);
    }
  }
}

// ---- reduced precision, fused with the row sums ------------------------------
//
// The coefficients enter the MFMAs of every angle wavefront of a point; they
// are converted once per call into 16 bytes per coefficient by
// k_s12_split_coef, in the index order of a_n, b_n (coef_off + k):
//   FP32   : float4 {U.re, V.re, U.im, V.im}  (float2 (U_c, V_c) at 2 c + comp)
//   BF16X3 : uint2 per component {U_hi | U_lo << 16, V_hi | V_lo << 16}
// with U = a + b, V = a - b formed in FP64 and split as k_s12_bf16x3 does, so
// the products are bit for bit those of the unfused kernels.
__global__ void k_s12_split_coef(int64_t n_coef, const double* __restrict__ an,
                                 const double* __restrict__ bn, int precision,
                                 uint4* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_coef) return;
  const double ar = an[2 * c], ai = an[2 * c + 1], br = bn[2 * c], bi = bn[2 * c + 1];
  const double ur = ar + br, ui = ai + bi, vr = ar - br, vi = ai - bi;
  uint4 o;
  if (precision == GM_PREC_FP32) {
    o.x = __float_as_uint((float)ur);
    o.y = __float_as_uint((float)vr);
    o.z = __float_as_uint((float)ui);
    o.w = __float_as_uint((float)vi);
  } else {
    unsigned short h, l;
    bf16_split(ur, h, l);
    o.x = h | ((unsigned)l << 16);
    bf16_split(vr, h, l);
    o.y = h | ((unsigned)l << 16);
    bf16_split(ui, h, l);
    o.z = h | ((unsigned)l << 16);
    bf16_split(vi, h, l);
    o.w = h | ((unsigned)l << 16);
  }
  out[c] = o;
}

// Epilogue of the reduced-precision row kernels: D[4 kq + r][i] of tile mt
// holds, in elements (0, 1), point p0 + 8 mt + 2 kq (re, im) and in (2, 3)
// the next; Mueller elements in FP64, the row sums over the chunk's points in
// point order, then over the lane groups, into the chunk's partials.
template <int MT, int RM>
__device__ __forceinline__ void rows_lp_epilogue(const float4_ (&accp)[MT][4],
                                                 const float4_ (&accm)[MT][4], int lane,
                                                 int64_t chunk, int j0, int n_ang, int64_t p0,
                                                 int64_t cend, int nr, const double* wr,
                                                 int64_t ldw, double* __restrict__ part) {
  const int i = lane & 15;
  const int kq = lane >> 4;
  double* dst = part + (size_t)chunk * RM * 4 * n_ang;
  double wv[MT][2][RM];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t p = p0 + 8 * mt + 2 * kq + h;
#pragma unroll
      for (int r = 0; r < RM; ++r)
        wv[mt][h][r] = (r < nr && p < cend) ? wr[(size_t)r * ldw + p] : 0.0;
    }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    double el[MT][2][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const cd sp = cd{(double)accp[mt][t][2 * h], (double)accp[mt][t][2 * h + 1]};
        const cd sm = cd{(double)accm[mt][t][2 * h], (double)accm[mt][t][2 * h + 1]};
        const cd s1 = (sp + sm) * 0.5, s2 = (sp - sm) * 0.5;
        const double a1 = norm2(s1), a2 = norm2(s2);
        const cd c = s1 * conj(s2);
        el[mt][h][0] = 0.5 * (a1 + a2);
        el[mt][h][1] = 0.5 * (a2 - a1);
        el[mt][h][2] = c.re;
        el[mt][h][3] = c.im;
      }
    const int j = j0 + t * 16 + i;
#pragma unroll
    for (int r = 0; r < RM; ++r) {
      double mine = 0.0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double v = 0.0;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int h = 0; h < 2; ++h) v = fma(wv[mt][h][r], el[mt][h][e], v);
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (e == kq) mine = v;
      }
      if (r < nr && j < n_ang) dst[((size_t)r * 4 + kq) * n_ang + j] = mine;
    }
  }
}

// S1/S2 of 8 MT points x 64 angles per wavefront on the FP32 or split-BF16
// matrix cores (operands as in k_s12_f32 / k_s12_bf16x3, coefficients from
// k_s12_split_coef), then the Mueller elements and their weighted row sums
// in FP64 as k_s12_rows_f64 forms them; same partial buffer and reduction
// (k_s12_rows_reduce).  A chunk is one group of MT 8-point tiles: each
// angle-table word fetched feeds the products of all MT tiles, and the row
// sums are formed once, after the K loop, so they hold no registers while
// it runs.
template <int PREC, int MT, int RM>
__global__ __launch_bounds__(64 * kRowsWaves) void k_s12_rows_lp(
    RowGroups g, const int64_t* __restrict__ off, const int32_t* __restrict__ nmax,
    const uint4* __restrict__ coef, int n_ang, int apad, int kpad, const void* __restrict__ tab,
    int nmax_cap, const double* __restrict__ w, int64_t ldw, double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t chunk = blockIdx.x;
  const int j0 = (blockIdx.y * kRowsWaves + (threadIdx.x >> 6)) * 64;
  if (j0 >= n_ang) return;   // whole wavefront past the last angle (no barriers below)
  int gi = 0;
  while (gi + 1 < g.n_groups && g.chunk0[gi + 1] <= chunk) ++gi;
  const int64_t p0 = g.start[gi] + (chunk - g.chunk0[gi]) * 8 * MT;
  const int64_t cend = min(p0 + 8 * MT, g.start[gi] + g.count[gi]);
  const int nr = g.nrow[gi];
  const double* wr = w + (size_t)g.row0[gi] * ldw;
  const int i = lane & 15;
  const int kq = lane >> 4;
  const int comp = i & 1;   // A row i of tile mt: point p0 + 8 mt + (i >> 1), component i & 1
  constexpr int kStep = PREC == GM_PREC_FP32 ? 4 : kLpK;
  int nm_a[MT];
  int64_t oa[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int64_t pa = p0 + 8 * mt + (i >> 1);
    const bool ok = pa < cend;
    nm_a[mt] = ok ? nmax[pa] : 0;
    oa[mt] = ok ? off[pa] : 0;
  }
  int kmax = 0;
  for (int q = 0; q < 8 * MT; ++q)
    if (p0 + q < cend) kmax = max(kmax, nmax[p0 + q]);
  kmax = (min(kmax, nmax_cap) + kStep - 1) / kStep * kStep;   // <= kpad
  float4_ accp[MT][4], accm[MT][4];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      accp[mt][t] = (float4_){0.f, 0.f, 0.f, 0.f};
      accm[mt][t] = (float4_){0.f, 0.f, 0.f, 0.f};
    }
  if constexpr (PREC == GM_PREC_FP32) {
    const float2* cf = (const float2*)coef;
    const float* Pt = (const float*)tab;
    const float* Mt = Pt + (size_t)kpad * apad;
    float2 uv_n[MT];
    float p_n[4], m_n[4];
    auto load_step = [&](int k0) {
      const int kk = k0 + kq;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
        uv_n[mt] = kk < nm_a[mt] ? cf[2 * (oa[mt] + kk) + comp] : make_float2(0.f, 0.f);
      const size_t row = (size_t)kk * apad + j0 + i;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        p_n[t] = Pt[row + 16 * t];
        m_n[t] = Mt[row + 16 * t];
      }
    };
    if (kmax > 0) load_step(0);
    for (int k0 = 0; k0 < kmax; k0 += 4) {
      float ua[MT], va[MT], P[4], M[4];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        ua[mt] = uv_n[mt].x;
        va[mt] = uv_n[mt].y;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        P[t] = p_n[t];
        M[t] = m_n[t];
      }
      if (k0 + 4 < kmax) load_step(k0 + 4);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          accp[mt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ua[mt], P[t], accp[mt][t], 0, 0, 0);
          accm[mt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(va[mt], M[t], accm[mt][t], 0, 0, 0);
        }
    }
  } else {
    const uint2* cb = (const uint2*)coef;
    const uint4* tb = (const uint4*)tab;   // 8 BF16 (one K-octet) per uint4
    const size_t plane4 = (size_t)kpad * apad / 8;
    // next step's operands: each 16-angle sub-tile's table words are reloaded
    // as soon as that sub-tile's products are issued, into the same registers
    uint4 tn[4][4];
    uint2 cn[MT][8];
    auto load_coef = [&](int k0) {
      const int kb = k0 + 8 * kq;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < 8; ++t)
          cn[mt][t] =
              kb + t < nm_a[mt] ? cb[2 * (oa[mt] + kb + t) + comp] : make_uint2(0u, 0u);
    };
    auto load_tab = [&](int k0, int t) {
      const size_t o4 = (size_t)((k0 + 8 * kq) >> 3) * apad + j0 + i + 16 * t;
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) tn[t][pl] = tb[pl * plane4 + o4];
    };
    // the operands of the current step, unpacked from the coefficient words
    // at the end of the step before (the words themselves live only
    // within a step: no copies of them cross the loop)
    union { bf16x8_ v; unsigned u[4]; } uh[MT], ul[MT], vh[MT], vl[MT];
    auto unpack = [&]() {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const uint2 a = cn[mt][2 * t], b = cn[mt][2 * t + 1];
          uh[mt].u[t] = (a.x & 0xffffu) | (b.x << 16);
          ul[mt].u[t] = (a.x >> 16) | (b.x & 0xffff0000u);
          vh[mt].u[t] = (a.y & 0xffffu) | (b.y << 16);
          vl[mt].u[t] = (a.y >> 16) | (b.y & 0xffff0000u);
        }
    };
    if (kmax > 0) {
      load_coef(0);
#pragma unroll
      for (int t = 0; t < 4; ++t) load_tab(0, t);
      unpack();
    }
    for (int k0 = 0; k0 < kmax; k0 += kLpK) {
      const bool more = k0 + kLpK < kmax;
      if (more) load_coef(k0 + kLpK);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8_ ph = __builtin_bit_cast(bf16x8_, tn[t][0]);
        const bf16x8_ pl = __builtin_bit_cast(bf16x8_, tn[t][1]);
        const bf16x8_ mh = __builtin_bit_cast(bf16x8_, tn[t][2]);
        const bf16x8_ ml = __builtin_bit_cast(bf16x8_, tn[t][3]);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          accp[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uh[mt].v, ph, accp[mt][t], 0, 0, 0);
          accp[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uh[mt].v, pl, accp[mt][t], 0, 0, 0);
          accp[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ul[mt].v, ph, accp[mt][t], 0, 0, 0);
          accm[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[mt].v, mh, accm[mt][t], 0, 0, 0);
          accm[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[mt].v, ml, accm[mt][t], 0, 0, 0);
          accm[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vl[mt].v, mh, accm[mt][t], 0, 0, 0);
        }
        // the sub-tile's words for the next step go into the registers its
        // products have just read (no copies of the buffer)
        if (more) load_tab(k0 + kLpK, t);
      }
      if (more) unpack();
    }
  }
  rows_lp_epilogue<MT, RM>(accp, accm, lane, chunk, j0, n_ang, p0, cend, nr, wr, ldw, part);
}

// Split-BF16 S1/S2 fused with the row sums, kShChunks chunks per block
// sharing each K-step's angle-table tile through LDS.  Wavefront w of block
// (bx, by) takes chunk kShChunks bx + w (8 MT points, as k_s12_rows_lp) and
// the 64 angles from 64 by.  The block's threads stage the next tile of kShK
// table rows (4 planes x 32 K x 64 angles, 16 KB) into the other half of a
// double buffer while the wavefronts issue the current tile's MFMAs, so each
// table word crosses the L2 once per kShChunks chunks and no wavefront holds
// table words of a later step in registers beyond its own staging share.
// Used with one 8-point tile per chunk, where it is the faster layout (the
// register-prefetching k_s12_rows_lp wins with two).  The block's K loop runs to the
// largest series length of its chunks; a wavefront whose own chunk is done
// still stages its share of the tiles.  Products, sums and the partial
// layout are those of k_s12_rows_lp (bit for bit: same operands, same
// accumulation order).
constexpr int kShChunks = 4;
constexpr int kShK = 32;   // K rows per staged tile (one MFMA step per barrier)
template <int MT, int RM>
__global__ __launch_bounds__(64 * kShChunks) void k_s12_rows_bf16s(
    RowGroups g, int64_t chunks, const int64_t* __restrict__ off,
    const int32_t* __restrict__ nmax, const uint4* __restrict__ coef, int n_ang, int apad,
    int kpad, const void* __restrict__ tab, int nmax_cap, const double* __restrict__ w,
    int64_t ldw, double* __restrict__ part) {
  // [buffer][plane][K-octet][angle], 2 x 16 KB
  __shared__ uint4 st[2][4][kShK / 8][64];
  __shared__ int s_kmax[kShChunks];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t chunk = (int64_t)blockIdx.x * kShChunks + wave;
  const bool live = chunk < chunks;
  const int j0 = blockIdx.y * 64;
  const int i = lane & 15;
  const int kq = lane >> 4;
  const int comp = i & 1;
  int gi = 0;
  int64_t p0 = 0, cend = 0;
  int nr = 0;
  const double* wr = w;
  int nm_a[MT];
  int64_t oa[MT];
  int kmax = 0;
  if (live) {
    while (gi + 1 < g.n_groups && g.chunk0[gi + 1] <= chunk) ++gi;
    p0 = g.start[gi] + (chunk - g.chunk0[gi]) * 8 * MT;
    cend = min(p0 + 8 * MT, g.start[gi] + g.count[gi]);
    nr = g.nrow[gi];
    wr = w + (size_t)g.row0[gi] * ldw;
    for (int q = 0; q < 8 * MT; ++q)
      if (p0 + q < cend) kmax = max(kmax, nmax[p0 + q]);
    kmax = (min(kmax, nmax_cap) + kLpK - 1) / kLpK * kLpK;   // <= kpad
  }
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int64_t pa = p0 + 8 * mt + (i >> 1);
    const bool ok = live && pa < cend;
    nm_a[mt] = ok ? nmax[pa] : 0;
    oa[mt] = ok ? off[pa] : 0;
  }
  if (lane == 0) s_kmax[wave] = kmax;
  __syncthreads();
  int kmax_b = 0;
#pragma unroll
  for (int q = 0; q < kShChunks; ++q) kmax_b = max(kmax_b, s_kmax[q]);
  float4_ accp[MT][4], accm[MT][4];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      accp[mt][t] = (float4_){0.f, 0.f, 0.f, 0.f};
      accm[mt][t] = (float4_){0.f, 0.f, 0.f, 0.f};
    }
  const uint2* cb = (const uint2*)coef;
  const uint4* tb = (const uint4*)tab;   // 8 BF16 (one K-octet) per uint4
  const size_t plane4 = (size_t)kpad * apad / 8;
  // this thread's staging words: K-octet tid / 64 of the tile, angle tid % 64,
  // every plane (kpad is a multiple of kShK and apad of 64: the tile lies
  // inside the zero-padded table)
  const int so8 = tid >> 6, sa = tid & 63;
  const size_t so = (size_t)so8 * apad + j0 + sa;
  uint4 g0, g1, g2, g3;   // one word per plane
#define GM_STAGE_LOAD(k0_)                                       \
  {                                                              \
    const size_t o_ = (size_t)((k0_) >> 3) * apad + so;          \
    g0 = tb[o_];                                                 \
    g1 = tb[plane4 + o_];                                        \
    g2 = tb[2 * plane4 + o_];                                    \
    g3 = tb[3 * plane4 + o_];                                    \
  }
#define GM_STAGE_STORE(b_)                                       \
  {                                                              \
    st[b_][0][so8][sa] = g0;                                     \
    st[b_][1][so8][sa] = g1;                                     \
    st[b_][2][so8][sa] = g2;                                     \
    st[b_][3][so8][sa] = g3;                                     \
  }
  uint2 cn[MT][8];
  auto load_coef = [&](int k0) {
    const int kb = k0 + 8 * kq;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int t = 0; t < 8; ++t)
        cn[mt][t] = kb + t < nm_a[mt] ? cb[2 * (oa[mt] + kb + t) + comp] : make_uint2(0u, 0u);
  };
  if (kmax_b > 0) {
    GM_STAGE_LOAD(0);
    GM_STAGE_STORE(0);
    load_coef(0);
  }
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < kmax_b; k0 += kShK) {
    const bool more = k0 + kShK < kmax_b;
    if (more) GM_STAGE_LOAD(k0 + kShK);
#pragma unroll
    for (int sub = 0; sub < kShK / kLpK; ++sub) {
      const int ks = k0 + sub * kLpK;
      if (ks < kmax) {   // this wavefront's own chunk still has terms
        union { bf16x8_ v; unsigned u[4]; } uh[MT], ul[MT], vh[MT], vl[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const uint2 a = cn[mt][2 * t], b = cn[mt][2 * t + 1];
            uh[mt].u[t] = (a.x & 0xffffu) | (b.x << 16);
            ul[mt].u[t] = (a.x >> 16) | (b.x & 0xffff0000u);
            vh[mt].u[t] = (a.y & 0xffffu) | (b.y << 16);
            vl[mt].u[t] = (a.y >> 16) | (b.y & 0xffff0000u);
          }
        if (ks + kLpK < kmax) load_coef(ks + kLpK);
        const int ko = 4 * sub + kq;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          union { uint4 q; bf16x8_ v; } ph, pl, mh, ml;
          ph.q = st[buf][0][ko][16 * t + i];
          pl.q = st[buf][1][ko][16 * t + i];
          mh.q = st[buf][2][ko][16 * t + i];
          ml.q = st[buf][3][ko][16 * t + i];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) {
            accp[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uh[mt].v, ph.v, accp[mt][t], 0, 0, 0);
            accp[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uh[mt].v, pl.v, accp[mt][t], 0, 0, 0);
            accp[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ul[mt].v, ph.v, accp[mt][t], 0, 0, 0);
            accm[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[mt].v, mh.v, accm[mt][t], 0, 0, 0);
            accm[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[mt].v, ml.v, accm[mt][t], 0, 0, 0);
            accm[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vl[mt].v, mh.v, accm[mt][t], 0, 0, 0);
          }
        }
      }
    }
    // the other buffer was last read in the previous tile, before the barrier
    if (more) GM_STAGE_STORE(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#undef GM_STAGE_STORE
#undef GM_STAGE_LOAD
  if (live)
    rows_lp_epilogue<MT, RM>(accp, accm, lane, chunk, j0, n_ang, p0, cend, nr, wr, ldw, part);
}

}  // namespace gm

namespace {

// Launch the S1/S2 contraction at the requested precision.  Tables for the
// reduced-precision paths are rebuilt on every call (one pass over pi, tau:
// small next to the contraction itself).
int launch_s12(gm_ctx* ctx, const char* who, int64_t n, const int64_t* off, const int32_t* nmax,
               const double* an, const double* bn, int n_ang, const double* pi_t,
               const double* tau_t, int nmax_cap, int mode, double* out0, double* out1,
               int precision, hipStream_t s) {
  if (precision != GM_PREC_FP64 && precision != GM_PREC_FP32 && precision != GM_PREC_BF16X3)
    return gm_fail(GM_EINVAL, "%s: unknown precision %d", who, precision);
  if (nmax_cap < 1) return gm_fail(GM_EINVAL, "%s: nmax_cap must be >= 1", who);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  const int wpb = ctx->tune_s12_waves;
  dim3 grid((unsigned)((n + 7) / 8), (unsigned)((n_ang + 64 * wpb - 1) / (64 * wpb)));
  dim3 block(64 * wpb);
  const int wave_points = ctx->tune_s12_layout;   // FP64 kernel only
  size_t slot;
  if (precision == GM_PREC_FP64) {
    ctx->mark_begin(gm::kKS12, s, &slot);
    const int pts = ctx->tune_s12_points == 16 ? 16 : 8;
    if (wave_points == 2) {   // LDS-staged angle tables, 8 or 16 points per wavefront
      const int64_t per_block = (int64_t)pts * gm::kS12LdsWaves;
      dim3 gl((unsigned)((n + per_block - 1) / per_block), (unsigned)((n_ang + 63) / 64));
      if (pts == 16)
        hipLaunchKernelGGL(gm::k_s12_f64_lds<2>, gl, dim3(64 * gm::kS12LdsWaves), 0, s, n, off,
                           nmax, an, bn, n_ang, pi_t, tau_t, nmax_cap, mode, out0, out1);
      else
        hipLaunchKernelGGL(gm::k_s12_f64_lds<1>, gl, dim3(64 * gm::kS12LdsWaves), 0, s, n, off,
                           nmax, an, bn, n_ang, pi_t, tau_t, nmax_cap, mode, out0, out1);
      ctx->mark_end(slot, s);
      e = hipGetLastError();
      if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "%s: %s", who, hipGetErrorString(e));
      return GM_OK;
    }
    dim3 g = wave_points
                 ? dim3((unsigned)((n + (int64_t)pts * wpb - 1) / ((int64_t)pts * wpb)),
                        (unsigned)((n_ang + 63) / 64))
                 : dim3((unsigned)((n + pts - 1) / pts), grid.y);
    if (pts == 16)
      hipLaunchKernelGGL(gm::k_s12_f64<2>, g, block, 0, s, n, off, nmax, an, bn, n_ang, pi_t,
                         tau_t, nmax_cap, mode, wave_points, out0, out1);
    else
      hipLaunchKernelGGL(gm::k_s12_f64<1>, g, block, 0, s, n, off, nmax, an, bn, n_ang, pi_t,
                         tau_t, nmax_cap, mode, wave_points, out0, out1);
    ctx->mark_end(slot, s);
  } else {
    const int kpad = (nmax_cap + gm::kLpK - 1) / gm::kLpK * gm::kLpK;
    // whole 256-angle blocks: every wavefront with j0 < n_ang reads 64 padded columns
    const int apad = (n_ang + 255) / 256 * 256;
    const size_t elems = (size_t)kpad * apad;
    const size_t bytes = precision == GM_PREC_FP32 ? 2 * elems * sizeof(float)
                                                   : 4 * elems * sizeof(unsigned short);
    int rc = ctx->reserve(ctx->tab_lp, bytes);
    if (rc) return rc;
    float* t32 = (float*)ctx->tab_lp.ptr;
    unsigned short* t16 = (unsigned short*)ctx->tab_lp.ptr;
    hipLaunchKernelGGL(gm::k_s12_prep_tables, dim3((unsigned)((elems + 255) / 256)), dim3(256), 0,
                       s, n_ang, nmax_cap, kpad, apad, pi_t, tau_t, precision, t32, t16);
    ctx->mark_begin(gm::kKS12, s, &slot);
    if (precision == GM_PREC_FP32)
      hipLaunchKernelGGL(gm::k_s12_f32, grid, block, 0, s, n, off, nmax, an, bn, n_ang, apad,
                         kpad, (const float*)t32, nmax_cap, mode, out0, out1);
    else
      hipLaunchKernelGGL(gm::k_s12_bf16x3, grid, block, 0, s, n, off, nmax, an, bn, n_ang,
                         apad, kpad, (const unsigned short*)t16, nmax_cap, mode, out0, out1);
    ctx->mark_end(slot, s);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "%s: %s", who, hipGetErrorString(e));
  return GM_OK;
}

}  // namespace

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
  return launch_s12(ctx, "gm_mie_s12", n, coef_off, nmax, an, bn, n_ang, pi_t, tau_t, nmax_cap,
                    (int)gm::kOutS12, s1, s2, precision, (hipStream_t)stream);
}

int gm_s12_mueller(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
                   const double* an, const double* bn, int n_ang, const double* pi_t,
                   const double* tau_t, int nmax_cap, double* mueller, int precision,
                   gm_stream stream) {
  if (!ctx || n < 0 || n_ang < 0) return gm_fail(-1, "gm_s12_mueller: bad arguments");
  if (n == 0 || n_ang == 0) return 0;
  if (!coef_off || !nmax || !an || !bn || !pi_t || !tau_t || !mueller)
    return gm_fail(-1, "gm_s12_mueller: NULL array");
  return launch_s12(ctx, "gm_s12_mueller", n, coef_off, nmax, an, bn, n_ang, pi_t, tau_t,
                    nmax_cap, (int)gm::kOutMueller, mueller, (double*)nullptr, precision,
                    (hipStream_t)stream);
}

int gm_s12_split_coef(gm_ctx* ctx, int64_t n_coef, const double* an, const double* bn,
                      int precision, void* coef_lp, gm_stream stream) {
  if (!ctx || n_coef < 0 || (precision != GM_PREC_FP32 && precision != GM_PREC_BF16X3))
    return gm_fail(GM_EINVAL, "gm_s12_split_coef: bad arguments (precision FP32 or BF16X3)");
  if (n_coef == 0) return GM_OK;
  if (!an || !bn || !coef_lp) return gm_fail(GM_EINVAL, "gm_s12_split_coef: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  size_t slot;
  ctx->mark_begin(gm::kKS12Prep, (hipStream_t)stream, &slot);
  hipLaunchKernelGGL(gm::k_s12_split_coef, dim3((unsigned)((n_coef + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n_coef, an, bn, precision, (uint4*)coef_lp);
  ctx->mark_end(slot, (hipStream_t)stream);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "gm_s12_split_coef: %s", hipGetErrorString(e));
  return GM_OK;
}

int gm_s12_contract_rows(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
                         const double* an, const double* bn, int n_ang, const double* pi_t,
                         const double* tau_t, int nmax_cap, int n_rows, const int64_t* start_host,
                         const int64_t* count_host, const double* w, double* out, int precision,
                         const void* coef_lp, gm_stream stream) {
  if (!ctx || n < 0 || n_ang < 0 || n_rows < 0 || n_rows > gm::kRowsPerCall)
    return gm_fail(GM_EINVAL, "gm_s12_contract_rows: bad arguments (at most %d rows)",
                   gm::kRowsPerCall);
  if (precision != GM_PREC_FP64 && precision != GM_PREC_FP32 && precision != GM_PREC_BF16X3)
    return gm_fail(GM_EINVAL, "gm_s12_contract_rows: unknown precision %d", precision);
  if (n_rows == 0 || n_ang == 0) return GM_OK;
  if (!start_host || !count_host || !coef_off || !nmax || !pi_t || !tau_t || !w || !out)
    return gm_fail(GM_EINVAL, "gm_s12_contract_rows: NULL array");
  if (precision == GM_PREC_FP64 ? (!an || !bn) : !coef_lp)
    return gm_fail(GM_EINVAL, precision == GM_PREC_FP64
                                  ? "gm_s12_contract_rows: NULL a_n / b_n"
                                  : "gm_s12_contract_rows: reduced precision needs the "
                                    "coefficients of gm_s12_split_coef (coef_lp)");
  if (nmax_cap < 1) return gm_fail(GM_EINVAL, "gm_s12_contract_rows: nmax_cap must be >= 1");
  // rows with the same range form a group (at most 4 rows; a longer run of
  // equal ranges is split, its points then contracted once per group).  One
  // 8-point tile per chunk, two when that still leaves >= 16k wavefronts.
  int64_t tiles = 0;
  for (int r = 0; r < n_rows; ++r)
    if (count_host[r] > 0) tiles += (count_host[r] + 7) / 8;
  const int64_t ablocks = (n_ang + 64 * gm::kRowsWaves - 1) / (64 * gm::kRowsWaves);
  // most rows of one group (consecutive rows with the same range, at most 4)
  int rm_pre = 1;
  for (int r = 1, run = 1; r < n_rows; ++r) {
    run = (start_host[r] == start_host[r - 1] && count_host[r] == count_host[r - 1] && run < 4)
              ? run + 1 : 1;
    rm_pre = max(rm_pre, run);
  }
  // tiles per chunk.  The kernels multiply two tiles together, sharing each
  // angle-table fetch (FP64 with one row sum per group; otherwise FP64 walks
  // the tiles one after the other, which pays for short series only: fewer
  // work units and half the partial sums where the K loop is short).  Two
  // when that still leaves >= 16k wavefronts at FP64, >= 8k at FP32 and
  // split-BF16.
  const int64_t waves2 = (tiles / 2) * ablocks * gm::kRowsWaves;
  const int row_tiles =
      ctx->tune_s12_row_tiles
          ? ctx->tune_s12_row_tiles
          : (precision == GM_PREC_FP64
                 ? ((rm_pre == 1 || nmax_cap <= 256) && waves2 >= 16384 ? 2 : 1)
                 : (waves2 >= 8192 ? 2 : 1));
  const int chunk_pts = 8 * row_tiles;
  gm::RowGroups g;
  g.n_groups = 0;
  g.chunk0[0] = 0;
  int rm = 1;
  for (int r = 0; r < n_rows; ++r) {
    if (start_host[r] < 0 || count_host[r] < 1 || start_host[r] + count_host[r] > n)
      return gm_fail(GM_EINVAL, "gm_s12_contract_rows: row %d range outside [0, n)", r);
    const int k = g.n_groups - 1;
    if (k >= 0 && start_host[r] == g.start[k] && count_host[r] == g.count[k] && g.nrow[k] < 4) {
      rm = max(rm, ++g.nrow[k]);
      continue;
    }
    g.start[g.n_groups] = start_host[r];
    g.count[g.n_groups] = count_host[r];
    g.row0[g.n_groups] = r;
    g.nrow[g.n_groups] = 1;
    g.chunk0[g.n_groups + 1] = g.chunk0[g.n_groups] + (count_host[r] + chunk_pts - 1) / chunk_pts;
    ++g.n_groups;
  }
  for (int k = g.n_groups; k < gm::kRowsPerCall; ++k) {
    g.start[k] = g.count[k] = 0;
    g.row0[k] = n_rows;
    g.nrow[k] = 0;
    g.chunk0[k + 1] = g.chunk0[g.n_groups];
  }
  const int64_t chunks = g.chunk0[g.n_groups];
  if (chunks > 0x7fffffffLL) return gm_fail(GM_EINVAL, "gm_s12_contract_rows: too many points");
  const int RM = rm <= 1 ? 1 : (rm <= 2 ? 2 : 4);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  int rc = ctx->reserve(ctx->gemm_tmp, sizeof(double) * (size_t)chunks * RM * 4 * n_ang);
  if (rc) return rc;
  double* part = (double*)ctx->gemm_tmp.ptr;
  const dim3 grid((unsigned)chunks, (unsigned)ablocks);
  const dim3 block(64 * gm::kRowsWaves);
  const dim3 rgrid((unsigned)((int64_t)n_rows * ((4 * (int64_t)n_ang + 63) / 64)));
  const dim3 rblock(64 * gm::kReduceWaves);
  size_t slot;
  if (precision == GM_PREC_FP64) {
    ctx->mark_begin(gm::kKS12, s, &slot);
    // two tiles multiplied together where a chunk has two and one row sum per
    // group (one table load feeds both tiles; with more row sums two tiles
    // leave one wavefront per SIMD)
#define GM_ROWS_F64(R_, M_)                                                                    \
  hipLaunchKernelGGL((gm::k_s12_rows_f64<R_, M_>), grid, block, 0, s, g, chunk_pts, coef_off, \
                     nmax, an, bn, n_ang, pi_t, tau_t, nmax_cap, w, n, part)
    if (RM == 1) {
      if (row_tiles >= 2) GM_ROWS_F64(1, 2); else GM_ROWS_F64(1, 1);
    } else if (RM == 2) {
      GM_ROWS_F64(2, 1);
    } else {
      GM_ROWS_F64(4, 1);
    }
#undef GM_ROWS_F64
    ctx->mark_end(slot, s);
  } else {
    // P, M tables of this precision (padded to whole K-steps and 256 angles)
    const int kpad = (nmax_cap + gm::kLpK - 1) / gm::kLpK * gm::kLpK;
    const int apad = (n_ang + 255) / 256 * 256;
    const size_t elems = (size_t)kpad * apad;
    const size_t bytes = precision == GM_PREC_FP32 ? 2 * elems * sizeof(float)
                                                   : 4 * elems * sizeof(unsigned short);
    rc = ctx->reserve(ctx->tab_lp, bytes);
    if (rc) return rc;
    ctx->mark_begin(gm::kKS12, s, &slot);
    hipLaunchKernelGGL(gm::k_s12_prep_tables, dim3((unsigned)((elems + 255) / 256)), dim3(256), 0,
                       s, n_ang, nmax_cap, kpad, apad, pi_t, tau_t, precision,
                       (float*)ctx->tab_lp.ptr, (unsigned short*)ctx->tab_lp.ptr);
    const uint4* cf = (const uint4*)coef_lp;
    const void* tb = ctx->tab_lp.ptr;
#define GM_ROWS_LP(P_, M_, R_)                                                                \
  hipLaunchKernelGGL((gm::k_s12_rows_lp<P_, M_, R_>), grid, block, 0, s, g, coef_off, nmax, cf,  \
                     n_ang, apad, kpad, tb, nmax_cap, w, n, part)
#define GM_ROWS_LP_RM(P_, M_)        \
  if (RM == 1) GM_ROWS_LP(P_, M_, 1); \
  else if (RM == 2) GM_ROWS_LP(P_, M_, 2); \
  else GM_ROWS_LP(P_, M_, 4)
    if (precision == GM_PREC_FP32) {
      if (row_tiles == 4) { GM_ROWS_LP_RM(GM_PREC_FP32, 4); }
      else if (row_tiles == 2) { GM_ROWS_LP_RM(GM_PREC_FP32, 2); }
      else { GM_ROWS_LP_RM(GM_PREC_FP32, 1); }
    } else if (ctx->tune_s12_share && row_tiles == 1) {
      // kShChunks work units per block share the angle-table tiles (LDS)
      const dim3 sgrid((unsigned)((chunks + gm::kShChunks - 1) / gm::kShChunks),
                       (unsigned)((n_ang + 63) / 64));
#define GM_ROWS_SH(M_, R_)                                                                     \
  hipLaunchKernelGGL((gm::k_s12_rows_bf16s<M_, R_>), sgrid, dim3(64 * gm::kShChunks), 0, s, g, \
                     chunks, coef_off, nmax, cf, n_ang, apad, kpad, tb, nmax_cap, w, n, part)
#define GM_ROWS_SH_RM(M_)        \
  if (RM == 1) GM_ROWS_SH(M_, 1); \
  else if (RM == 2) GM_ROWS_SH(M_, 2); \
  else GM_ROWS_SH(M_, 4)
      GM_ROWS_SH_RM(1);
#undef GM_ROWS_SH_RM
#undef GM_ROWS_SH
    } else {
      if (row_tiles == 4) { GM_ROWS_LP_RM(GM_PREC_BF16X3, 4); }
      else if (row_tiles == 2) { GM_ROWS_LP_RM(GM_PREC_BF16X3, 2); }
      else { GM_ROWS_LP_RM(GM_PREC_BF16X3, 1); }
    }
#undef GM_ROWS_LP_RM
#undef GM_ROWS_LP
    ctx->mark_end(slot, s);
  }
  ctx->mark_begin(gm::kKPsd, s, &slot);
  if (RM == 1)
    hipLaunchKernelGGL(gm::k_s12_rows_reduce<1>, rgrid, rblock, 0, s, g, n_ang,
                       (const double*)part, out);
  else if (RM == 2)
    hipLaunchKernelGGL(gm::k_s12_rows_reduce<2>, rgrid, rblock, 0, s, g, n_ang,
                       (const double*)part, out);
  else
    hipLaunchKernelGGL(gm::k_s12_rows_reduce<4>, rgrid, rblock, 0, s, g, n_ang,
                       (const double*)part, out);
  ctx->mark_end(slot, s);
  e = hipGetLastError();
  if (e != hipSuccess)
    return gm_fail(GM_ERUNTIME, "gm_s12_contract_rows: %s", hipGetErrorString(e));
  return GM_OK;
}

int gm_s12_p11(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
               const double* an, const double* bn, int n_ang, const double* pi_t,
               const double* tau_t, int nmax_cap, double* p11, int precision, gm_stream stream) {
  if (!ctx || n < 0 || n_ang < 0) return gm_fail(-1, "gm_s12_p11: bad arguments");
  if (n == 0 || n_ang == 0) return 0;
  if (!coef_off || !nmax || !an || !bn || !pi_t || !tau_t || !p11)
    return gm_fail(-1, "gm_s12_p11: NULL array");
  return launch_s12(ctx, "gm_s12_p11", n, coef_off, nmax, an, bn, n_ang, pi_t, tau_t, nmax_cap,
                    (int)gm::kOutP11, p11, (double*)nullptr, precision, (hipStream_t)stream);
}

}  // extern "C"
