// Context object behind the opaque gm_ctx handle of include/geosmie.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <initializer_list>
#include <vector>

#include "geosmie.h"

int gm_fail(int code, const char* fmt, ...);

struct Buf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct gm_ctx {
  int device = 0;
  int n_cu = 256;
  // resident Mie wavefronts per CU, per register-allocation variant (index
  // 0: unconstrained, 1: >= 3 waves per SIMD, 2: >= 4 waves per SIMD)
  int occ_homog[3] = {8, 8, 8}, occ_stable[3] = {8, 8, 8}, occ_literal = 4;
  // resident wavefronts per CU of the checkpointed coated kernel: K = 4
  // (unconstrained, >= 3 waves per SIMD) and K = 8
  int occ_ckpt4[2] = {8, 8}, occ_ckpt5 = 8, occ_ckpt8 = 5;
  int occ_coop = 8;             // the homogeneous kernel with P lanes per point
  int occ_homog_ckpt = 8;       // the checkpointed homogeneous kernel (K = 4)

  size_t scratch_budget = (size_t)3 << 30;       // bytes of recurrence scratch
  // launch shapes chosen by gm_ctx_tune (an autotuner's knobs)
  int tune_mie_grid_mult = 1;   // persistent Mie grid = multiple of resident wavefronts
  int tune_s12_waves = 4;       // wavefronts (64 angles each) per S1/S2 block
  int tune_s12_points = 8;      // points per wavefront of the FP64 S1/S2 kernel (8 or 16)
  int tune_psd_chunk = 128;     // points per split-K chunk of the few-PSD contraction
  int tune_mlp_wterms = 1;      // BF16 terms per MLP weight (1 = BF16 weights, 2 = hi + lo)
  int tune_mie_waves = 0;       // Mie register variant: 0 unconstrained, 3 or 4 waves per SIMD
  int tune_s12_layout = 0;      // FP64 S1/S2 block: waves split angles (0) or points (1)
  int tune_mie_ckpt = 4;        // coated kernel: rows per checkpoint (4, 8) or 0 = every row stored
  int tune_mie_lanes = 0;       // homogeneous kernel: lanes per point (1, 16, 32, 64; 0 = by batch)
  int psd_quirks = 1;           // integratePSD quirks reproduced (GM_PSD_QUIRK_* bits)
  int tune_s12_share = 1;        // split-BF16 fused S1/S2: angle-table tiles shared through LDS
  int tune_s12_row_tiles = 0;    // reduced-precision fused S1/S2: 8-point tiles per wave (0 = auto)
  Buf status;                                    // int32 flags + scan scratch
  Buf order, key, hist, scratch, scan_tmp;
  Buf tab_pi, tab_tau, gemm_tmp, feat_tmp, gsf_tab, tab_lp, lut_tmp, mlp_ws;
  // cached GSF tables: key = (ng, n_ang, checksum of the angle grid)
  int gsf_ng = 0, gsf_nang = 0;
  double gsf_key = 0.0;
  std::vector<Buf*> all_bufs() {
    return {&status, &order, &key, &hist, &scratch, &scan_tmp, &tab_pi, &tab_tau, &gemm_tmp,
            &feat_tmp, &gsf_tab, &tab_lp, &lut_tmp, &mlp_ws};
  }
  int reserve(Buf& b, size_t bytes);
  ~gm_ctx();

  // Optional per-kernel timing: event pairs recorded on the launch stream.
  bool timing = false;
  struct Pair {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Pair> pairs;
  std::vector<hipEvent_t> pool;
  hipEvent_t take_event();
  void mark_begin(int kernel, hipStream_t s, size_t* slot);
  void mark_end(size_t slot, hipStream_t s);
};

namespace gm {
// kernel ids for gm_ctx_kernel_time
enum KernelId { kKMieHomog = 0, kKMieCoated = 1, kKS12 = 2, kKPsd = 3, kKGsf = 4, kKMlp = 5,
                kKMieCoatedLit = 6, kKLut = 7, kKS12Prep = 8, kKCount = 9 };
}  // namespace gm

namespace gm {
// In-place inclusive prefix sum of n int64 values (device) on stream s.
int scan_inclusive_i64(gm_ctx* ctx, int64_t* data, int64_t n, hipStream_t s);
// Make ctx->gsf_tab hold the GSF tables G[4][n_ang][ng] (P00, P22, P2-2, P02
// with Gauss weights and the angle interpolation folded in) for this grid.
int gsf_tables(gm_ctx* ctx, int ng, int n_ang, const double* ang_deg_host, hipStream_t s);
__global__ void k_nmax_i32(int64_t n, const double* size, int32_t* out);
}  // namespace gm
