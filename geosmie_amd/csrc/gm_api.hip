// C ABI of libgeosmie_hip.so (declared in include/geosmie.h).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "geosmie.h"
#include "gm_ctx.h"
#include "gm_internal.h"

static thread_local char g_err[512] = "";

int gm_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_OK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return gm_fail(GM_ERUNTIME, "%s: %s", #expr, hipGetErrorString(e_));         \
  } while (0)

int gm_ctx::reserve(Buf& b, size_t bytes) {
  if (b.bytes >= bytes) return GM_OK;
  if (b.ptr) {
    // Work queued earlier on any stream may still read the old buffer.
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipFree: %s", hipGetErrorString(e));
  }
  size_t want = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    return gm_fail(GM_ENOMEM, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
  }
  b.bytes = want;
  return GM_OK;
}

gm_ctx::~gm_ctx() {
  hipSetDevice(device);
  for (Buf* b : all_bufs())
    if (b->ptr) hipFree(b->ptr);
  for (auto& p : pairs) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  for (auto e : pool) hipEventDestroy(e);
}

hipEvent_t gm_ctx::take_event() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void gm_ctx::mark_begin(int kernel, hipStream_t s, size_t* slot) {
  *slot = (size_t)-1;
  if (!timing) return;
  Pair p{kernel, take_event(), take_event()};
  if (!p.a || !p.b) return;
  hipEventRecord(p.a, s);
  pairs.push_back(p);
  *slot = pairs.size() - 1;
}

void gm_ctx::mark_end(size_t slot, hipStream_t s) {
  if (slot == (size_t)-1 || slot >= pairs.size()) return;
  hipEventRecord(pairs[slot].b, s);
}

namespace {

// Static scratch of gm_mie_batch_noalloc: one copy per GPU, held until the
// process ends.  Host code never uses the name as an address: its device
// address is taken once per device (hipGetSymbolAddress in gm_ctx_create,
// with that device current) and kept in g_noalloc[device].
constexpr size_t kNoAllocBytes = (size_t)128 << 20;
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc[kMaxDevices] = {};

// Take this device's address of the static scratch (the device is current).
int noalloc_init(int device) {
  if (device >= kMaxDevices) return GM_OK;   // such devices only lack the no-alloc entry
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc[device]) return GM_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return gm_fail(GM_ERUNTIME, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_noalloc[device] = p;
  return GM_OK;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  return g_noalloc[device];
}

// Buffers of one gm_mie_batch call.
struct MieWs {
  int64_t* order = nullptr;
  int32_t* key = nullptr;
  int64_t* hist = nullptr;     // histogram [nb], cursors [nb], class ranges [2 kClasses]
  int32_t* status = nullptr;
  double* slots = nullptr;     // recurrence slots of the persistent kernels
};

// Persistent grid of each Mie kernel: one wavefront per resident slot
// (occupancy measured per kernel at context creation) times the tuned
// multiple, capped by the waves of work and by the slots the budget holds.
struct MieGrids {
  std::pair<int64_t, size_t> homog, stable, literal;
  int lanes = 1;      // homogeneous kernel: lanes per point
  int lcap = 0;       // rows per lane segment (lanes > 1)
  int64_t small_grid = 0;   // grid of the 4-lane kernel for the short series (lanes > 1)
};

// Lanes per point of the homogeneous kernel: the tuned value, or by batch:
// one point per lane once the batch fills the GPU (64 points per wavefront,
// four per SIMD), otherwise P lanes per point, more for longer series.
int mie_lanes(const gm_ctx* ctx, int64_t n, int rows_cap) {
  if (ctx->tune_mie_lanes) return ctx->tune_mie_lanes;
  // one point per lane once the batch fills the GPU: four wavefronts per
  // SIMD, or one for short series (n_max <= 64), where the cooperative
  // kernels' scans cost more than the serial rows they split
  if (n >= (int64_t)64 * ctx->n_cu * 4 * 4) return 1;
  if (rows_cap <= 64 && n >= (int64_t)64 * ctx->n_cu * 4) return 1;
  return rows_cap <= 300 ? 16 : rows_cap <= 1200 ? 32 : 64;
}

// The lane-per-point homogeneous kernel in its checkpointed form: with the
// coated kernel's checkpoint knob at 4 or 8 (both use K = 4 here: the
// homogeneous rows are 3 doubles, and its step's registers leave room for
// more resident wavefronts than 8-row blocks of LDS would) and no forced
// register budget (that variant is built on the full-row kernel).
bool homog_ckpt(const gm_ctx* ctx) { return ctx->tune_mie_ckpt != 0 && ctx->tune_mie_waves == 0; }

MieGrids mie_grids(const gm_ctx* ctx, int64_t n, int rows_cap, size_t budget) {
  const int64_t waves = (n + 63) / 64;
  auto grid_for = [&](int occ, int row_doubles) {
    const size_t slot_bytes = (size_t)rows_cap * 64 * row_doubles * sizeof(double);
    int64_t g = std::min<int64_t>(
        waves, (int64_t)ctx->n_cu * std::max(occ, 1) * std::max(ctx->tune_mie_grid_mult, 1));
    g = std::min<int64_t>(g, std::max<int64_t>(1, (int64_t)(budget / slot_bytes)));
    return std::make_pair(g, (size_t)g * slot_bytes);
  };
  const int var = ctx->tune_mie_waves == 4 ? 2 : ctx->tune_mie_waves == 3 ? 1 : 0;
  MieGrids g;
  g.lanes = mie_lanes(ctx, n, rows_cap);
  if (g.lanes == 1 && homog_ckpt(ctx)) {
    // a checkpointed slot: 3 doubles per lane per K rows
    const size_t slot_bytes = (size_t)64 * gm::homog_ckpt_lane_doubles(rows_cap, 4) * sizeof(double);
    int64_t gg = std::min<int64_t>(waves, (int64_t)ctx->n_cu * std::max(ctx->occ_homog_ckpt, 1) *
                                              std::max(ctx->tune_mie_grid_mult, 1));
    gg = std::min<int64_t>(gg, std::max<int64_t>(1, (int64_t)(budget / slot_bytes)));
    g.homog = std::make_pair(gg, (size_t)gg * slot_bytes);
  } else if (g.lanes == 1) {
    g.homog = grid_for(ctx->occ_homog[var], gm::kHomogRowDoubles);
  } else {
    g.lcap = (rows_cap + g.lanes - 1) / g.lanes;
    const size_t slot_bytes = (size_t)g.lcap * 64 * gm::kHomogRowDoubles * sizeof(double);
    const int64_t per_wave = 64 / g.lanes;
    int64_t gg = std::min<int64_t>((n + per_wave - 1) / per_wave,
                                   (int64_t)ctx->n_cu * std::max(ctx->occ_coop, 1) *
                                       std::max(ctx->tune_mie_grid_mult, 1));
    gg = std::min<int64_t>(gg, std::max<int64_t>(1, (int64_t)(budget / slot_bytes)));
    // the 4-lane kernel of the short series runs after it on the same slots
    const size_t small_slot = (size_t)gm::kSmallLcap * 64 * gm::kHomogRowDoubles * sizeof(double);
    int64_t gsm = std::min<int64_t>((n + 15) / 16, (int64_t)ctx->n_cu * std::max(ctx->occ_coop, 1));
    gsm = std::max<int64_t>(1, std::min<int64_t>(gsm, (int64_t)(budget / small_slot)));
    g.small_grid = gsm;
    g.homog = std::make_pair(gg, std::max((size_t)gg * slot_bytes, (size_t)gsm * small_slot));
  }
  g.literal = grid_for(ctx->occ_literal, gm::kCoatedRowDoubles);
  const int ck = ctx->tune_mie_ckpt;
  if (ck == 0) {
    g.stable = grid_for(ctx->occ_stable[var], gm::kCoatedRowDoubles);
  } else {
    // a checkpointed slot: coated_ckpt_lane_doubles per lane (no per-row planes
    // beyond the start-correction ones)
    const size_t slot_bytes =
        (size_t)64 * gm::coated_ckpt_lane_doubles(rows_cap, ck) * sizeof(double);
    const int occ = ck == 8   ? ctx->occ_ckpt8
                    : ck == 5 ? ctx->occ_ckpt5
                              : ctx->occ_ckpt4[var ? 1 : 0];
    int64_t gg = std::min<int64_t>(
        waves, (int64_t)ctx->n_cu * std::max(occ, 1) * std::max(ctx->tune_mie_grid_mult, 1));
    gg = std::min<int64_t>(gg, std::max<int64_t>(1, (int64_t)(budget / slot_bytes)));
    g.stable = std::make_pair(gg, (size_t)gg * slot_bytes);
  }
  return g;
}

int mie_batch_check(const gm_ctx* ctx, int64_t n, const double* x, const double* y,
                    const double* eps, const double* eps2, double size_max,
                    const int64_t* coef_off, const double* an, const double* bn,
                    const char* who) {
  if (!ctx) return gm_fail(GM_EINVAL, "%s: ctx is NULL", who);
  if (n < 0) return gm_fail(GM_EINVAL, "%s: n < 0", who);
  if (n == 0) return GM_OK;
  if (!x || !eps) return gm_fail(GM_EINVAL, "Must specify x and either eps or m.");
  if ((y != nullptr) != (eps2 != nullptr))
    return gm_fail(GM_EINVAL, "Must specify both y and m2 for coated particles.");
  if (coef_off && (!an || !bn)) return gm_fail(GM_EINVAL, "%s: an/bn NULL", who);
  if (!(size_max >= 0.0) || !isfinite(size_max))
    return gm_fail(GM_EINVAL, "%s: size_max must be finite and >= 0", who);
  // the coated kernels address a wavefront's scratch slot with 32-bit byte
  // offsets (range-checked buffer loads): one slot must stay below 2 GiB
  if (y && (double)std::max(1, gm_nmax(size_max)) * 64.0 * gm::kCoatedRowDoubles *
                   sizeof(double) >= 2147483648.0)
    return gm_fail(GM_EINVAL, "%s: coated size parameter %.6g too large (series > 2 GiB of "
                              "scratch per wavefront)", who, size_max);
  return GM_OK;
}

// Plan (classify, sort by cost) and the three class kernels of gm_mie_batch.
int mie_batch_run(gm_ctx* ctx, int64_t n, const double* x, const double* y, const double* eps,
                  const double* mu, const double* eps2, int rows_cap, const int64_t* coef_off,
                  double* an, double* bn, double* props, int32_t* nmax_out, uint32_t flags,
                  const MieWs& ws, const MieGrids& g, hipStream_t s) {
  const int nb = gm::kClasses * gm::kBuckets;
  int64_t* hist = ws.hist;
  int64_t* cursor = hist + nb;
  int64_t* range = cursor + nb;
  int64_t* ab = range + 2 * gm::kClasses;   // homogeneous class: many-lane part, few-lane part
  unsigned gblk = (unsigned)((n + 255) / 256);
  // points per thread of the plan kernels: 1 up to ~256k points (blocks stay
  // plentiful), growing to kPlanPer for the largest batches (fewer LDS
  // histogram flushes)
  const int per = (int)std::min<int64_t>(gm::kPlanPer, std::max<int64_t>(1, n / (256 * 1024)));
  const unsigned pblk = (unsigned)((n + 256 * (int64_t)per - 1) / (256 * (int64_t)per));
  if (flags & GM_MIE_SORT) {
    hipLaunchKernelGGL(gm::k_zero_i64, dim3((nb + 255) / 256), dim3(256), 0, s, hist, nb);
    hipLaunchKernelGGL(gm::k_plan_keys, dim3(pblk), dim3(256), 0, s, n, x, y, eps, mu, eps2,
                       ws.key, nmax_out, hist, ws.status, per);
    hipLaunchKernelGGL(gm::k_plan_scan, dim3(1), dim3(256), 0, s, hist, cursor, range,
                       gm::kSmallCost >> gm::kBucketShift, ab);
    hipLaunchKernelGGL(gm::k_plan_scatter, dim3(pblk), dim3(256), 0, s, n,
                       (const int32_t*)ws.key, cursor, ws.order, per);
  } else {
    // input order: each class kernel walks every point and keeps its own
    hipLaunchKernelGGL(gm::k_plan_identity, dim3(gblk), dim3(256), 0, s, n, ws.order, range, ab);
    if (nmax_out) {
      hipLaunchKernelGGL(gm::k_nmax_i32, dim3(gblk), dim3(256), 0, s, n, y ? y : x, nmax_out);
    }
  }
  HIP_OK(hipGetLastError());
  const int dn_extra = (flags & GM_MIE_DN_CONVERGED) ? -1 : 0;
  const int var = ctx->tune_mie_waves == 4 ? 2 : ctx->tune_mie_waves == 3 ? 1 : 0;
  const int64_t* order = ws.order;
  size_t slot;
  ctx->mark_begin(gm::kKMieHomog, s, &slot);
  auto homog = var == 2 ? gm::k_mie_homog<4> : gm::k_mie_homog<1>;
  if (homog_ckpt(ctx)) homog = gm::k_mie_homog_ckpt<1, 4>;
  int cap_arg = rows_cap;
  const int64_t* hrange = range;
  if (g.lanes > 1) {
    homog = g.lanes == 16 ? gm::k_mie_homog_coop<16>
            : g.lanes == 32 ? gm::k_mie_homog_coop<32> : gm::k_mie_homog_coop<64>;
    cap_arg = g.lcap;
    hrange = ab;
  }
  hipLaunchKernelGGL(homog, dim3((unsigned)g.homog.first), dim3(64), 0, s, order, hrange, n, x, y,
                     eps, mu, eps2, coef_off, an, bn, props, ws.status, ws.slots, cap_arg,
                     dn_extra);
  if (g.lanes > 1 && (flags & GM_MIE_SORT)) {
    // the short series (sorted last): 4 lanes per point, 16 points per wavefront
    hipLaunchKernelGGL(gm::k_mie_homog_coop<4>, dim3((unsigned)g.small_grid), dim3(64), 0, s,
                       order, ab + 2, n, x, y, eps, mu, eps2, coef_off, an, bn, props, ws.status,
                       ws.slots, gm::kSmallLcap, dn_extra);
  }
  ctx->mark_end(slot, s);
  HIP_OK(hipGetLastError());
  if (y) {
    ctx->mark_begin(gm::kKMieCoated, s, &slot);
    auto stable = var == 2   ? gm::k_mie_coated_stable<4>
                  : var == 1 ? gm::k_mie_coated_stable<3>
                             : gm::k_mie_coated_stable<1>;
    if (ctx->tune_mie_ckpt == 4)
      stable = var ? gm::k_mie_coated_ckpt<3, 4> : gm::k_mie_coated_ckpt<1, 4>;
    else if (ctx->tune_mie_ckpt == 5)
      stable = gm::k_mie_coated_ckpt<1, 5>;
    else if (ctx->tune_mie_ckpt == 8)
      stable = gm::k_mie_coated_ckpt<1, 8>;
    hipLaunchKernelGGL(stable, dim3((unsigned)g.stable.first), dim3(64), 0, s, order, range, n, x,
                       y, eps, mu, eps2, coef_off, an, bn, props, ws.status, ws.slots, rows_cap,
                       dn_extra);
    ctx->mark_end(slot, s);
    HIP_OK(hipGetLastError());
    ctx->mark_begin(gm::kKMieCoatedLit, s, &slot);
    hipLaunchKernelGGL(gm::k_mie_coated_literal, dim3((unsigned)g.literal.first), dim3(64), 0, s,
                       order, range, n, x, y, eps, mu, eps2, coef_off, an, bn, props, ws.status,
                       ws.slots, rows_cap, dn_extra);
    ctx->mark_end(slot, s);
    HIP_OK(hipGetLastError());
  }
  return GM_OK;
}

}  // namespace

extern "C" {

const char* gm_version(void) { return "geosmie_amd 0.1 (gfx950)"; }
const char* gm_last_error(void) { return g_err; }

int gm_ctx_create(int device, gm_ctx** out) {
  if (!out) return gm_fail(GM_EINVAL, "gm_ctx_create: out is NULL");
  int ndev = 0;
  HIP_OK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return gm_fail(GM_EINVAL, "gm_ctx_create: device %d of %d", device, ndev);
  HIP_OK(hipSetDevice(device));
  gm_ctx* c = new gm_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete c;
    return gm_fail(GM_ERUNTIME, "hipGetDeviceProperties: %s", hipGetErrorString(e));
  }
  c->n_cu = prop.multiProcessorCount;
  // resident one-wavefront blocks per CU of each persistent Mie kernel
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_homog<1>, 64, 0) == hipSuccess)
    c->occ_homog[0] = c->occ_homog[1] = occ;   // the unconstrained homog kernel allows 3 already
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_homog<4>, 64, 0) == hipSuccess)
    c->occ_homog[2] = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_homog_ckpt<1, 4>, 64, 0) ==
      hipSuccess)
    c->occ_homog_ckpt = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_coated_stable<1>, 64, 0) ==
      hipSuccess)
    c->occ_stable[0] = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_coated_stable<3>, 64, 0) ==
      hipSuccess)
    c->occ_stable[1] = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_coated_stable<4>, 64, 0) ==
      hipSuccess)
    c->occ_stable[2] = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_coated_literal, 64, 0) ==
      hipSuccess)
    c->occ_literal = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_coated_ckpt<1, 4>, 64, 0) ==
      hipSuccess)
    c->occ_ckpt4[0] = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_coated_ckpt<3, 4>, 64, 0) ==
      hipSuccess)
    c->occ_ckpt4[1] = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_coated_ckpt<1, 5>, 64, 0) ==
      hipSuccess)
    c->occ_ckpt5 = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_coated_ckpt<1, 8>, 64, 0) ==
      hipSuccess)
    c->occ_ckpt8 = occ;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, gm::k_mie_homog_coop<16>, 64, 0) ==
      hipSuccess)
    c->occ_coop = occ;
  int rc = noalloc_init(device);
  if (!rc) rc = c->reserve(c->status, 64);
  if (rc) {
    delete c;
    return rc;
  }
  e = hipMemset(c->status.ptr, 0, 64);
  if (e != hipSuccess) {
    delete c;
    return gm_fail(GM_ERUNTIME, "hipMemset: %s", hipGetErrorString(e));
  }
  *out = c;
  return GM_OK;
}

int gm_ctx_destroy(gm_ctx* ctx) {
  if (!ctx) return GM_OK;
  delete ctx;
  return GM_OK;
}

int gm_ctx_check(gm_ctx* ctx, gm_stream stream, int32_t* flags_host) {
  if (!ctx || !flags_host) return gm_fail(GM_EINVAL, "gm_ctx_check: NULL argument");
  HIP_OK(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  int32_t v = 0;
  HIP_OK(hipMemcpyAsync(&v, ctx->status.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  HIP_OK(hipMemsetAsync(ctx->status.ptr, 0, sizeof(int32_t), s));
  HIP_OK(hipStreamSynchronize(s));
  *flags_host = v;
  return GM_OK;
}

int gm_ctx_tune(gm_ctx* ctx, int param, int value) {
  if (!ctx) return gm_fail(GM_EINVAL, "gm_ctx_tune: ctx is NULL");
  switch (param) {
    case GM_TUNE_MIE_GRID_MULT:
      if (value < 1 || value > 16) return gm_fail(GM_EINVAL, "gm_ctx_tune: grid multiple %d", value);
      ctx->tune_mie_grid_mult = value;
      return GM_OK;
    case GM_TUNE_S12_WAVES:
      if (value != 1 && value != 2 && value != 4)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: S1/S2 wavefronts per block must be 1, 2 or 4");
      ctx->tune_s12_waves = value;
      return GM_OK;
    case GM_TUNE_PSD_CHUNK:
      if (value < 16 || value > (1 << 20))
        return gm_fail(GM_EINVAL, "gm_ctx_tune: split-K chunk %d out of [16, 2^20]", value);
      ctx->tune_psd_chunk = value;
      return GM_OK;
    case GM_TUNE_S12_POINTS:
      if (value != 8 && value != 16)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: S1/S2 points per wavefront must be 8 or 16");
      ctx->tune_s12_points = value;
      return GM_OK;
    case GM_TUNE_S12_LAYOUT:
      if (value < 0 || value > 2)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: S1/S2 layout must be 0, 1 or 2");
      ctx->tune_s12_layout = value;
      return GM_OK;
    case GM_TUNE_MIE_WAVES:
      if (value != 0 && value != 3 && value != 4)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: Mie waves per SIMD must be 0, 3 or 4");
      ctx->tune_mie_waves = value;
      return GM_OK;
    case GM_TUNE_MIE_CKPT:
      if (value != 0 && value != 4 && value != 5 && value != 8)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: rows per checkpoint must be 0, 4, 5 or 8");
      ctx->tune_mie_ckpt = value;
      return GM_OK;
    case GM_TUNE_MIE_LANES:
      if (value != 0 && value != 1 && value != 16 && value != 32 && value != 64)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: lanes per point must be 0, 1, 16, 32 or 64");
      ctx->tune_mie_lanes = value;
      return GM_OK;
    case GM_TUNE_S12_ROW_TILES:
      if (value != 0 && value != 1 && value != 2 && value != 4)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: S1/S2 tiles per work unit must be 0, 1, 2 or 4");
      ctx->tune_s12_row_tiles = value;
      return GM_OK;
    case GM_TUNE_S12_SHARE:
      if (value != 0 && value != 1)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: S1/S2 table sharing must be 0 or 1");
      ctx->tune_s12_share = value;
      return GM_OK;
    case GM_TUNE_PSD_QUIRKS:
      if (value & ~(int)GM_PSD_QUIRK_ALIAS)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: unknown integratePSD quirk bits %d", value);
      ctx->psd_quirks = value;
      return GM_OK;
    case GM_TUNE_MLP_WEIGHT_TERMS:
      if (value != 1 && value != 2)
        return gm_fail(GM_EINVAL, "gm_ctx_tune: MLP weight terms must be 1 or 2");
      ctx->tune_mlp_wterms = value;
      return GM_OK;
    default:
      return gm_fail(GM_EINVAL, "gm_ctx_tune: unknown parameter %d", param);
  }
}

int gm_ctx_get_tune(gm_ctx* ctx, int param, int* value) {
  if (!ctx || !value) return gm_fail(GM_EINVAL, "gm_ctx_get_tune: NULL argument");
  switch (param) {
    case GM_TUNE_MIE_GRID_MULT: *value = ctx->tune_mie_grid_mult; return GM_OK;
    case GM_TUNE_S12_WAVES: *value = ctx->tune_s12_waves; return GM_OK;
    case GM_TUNE_PSD_CHUNK: *value = ctx->tune_psd_chunk; return GM_OK;
    case GM_TUNE_MLP_WEIGHT_TERMS: *value = ctx->tune_mlp_wterms; return GM_OK;
    case GM_TUNE_S12_POINTS: *value = ctx->tune_s12_points; return GM_OK;
    case GM_TUNE_MIE_WAVES: *value = ctx->tune_mie_waves; return GM_OK;
    case GM_TUNE_S12_LAYOUT: *value = ctx->tune_s12_layout; return GM_OK;
    case GM_TUNE_MIE_CKPT: *value = ctx->tune_mie_ckpt; return GM_OK;
    case GM_TUNE_MIE_LANES: *value = ctx->tune_mie_lanes; return GM_OK;
    case GM_TUNE_PSD_QUIRKS: *value = ctx->psd_quirks; return GM_OK;
    case GM_TUNE_S12_ROW_TILES: *value = ctx->tune_s12_row_tiles; return GM_OK;
    case GM_TUNE_S12_SHARE: *value = ctx->tune_s12_share; return GM_OK;
    default: return gm_fail(GM_EINVAL, "gm_ctx_get_tune: unknown parameter %d", param);
  }
}

int gm_ctx_set_timing(gm_ctx* ctx, int enable) {
  if (!ctx) return gm_fail(GM_EINVAL, "gm_ctx_set_timing: ctx is NULL");
  ctx->timing = enable != 0;
  return GM_OK;
}

int gm_ctx_kernel_time(gm_ctx* ctx, int kernel, double* ms_total, int64_t* launches) {
  if (!ctx || !ms_total || !launches || kernel < 0 || kernel >= gm::kKCount)
    return gm_fail(GM_EINVAL, "gm_ctx_kernel_time: bad arguments");
  HIP_OK(hipSetDevice(ctx->device));
  double tot = 0.0;
  int64_t cnt = 0;
  std::vector<gm_ctx::Pair> keep;
  for (auto& p : ctx->pairs) {
    if (p.kernel != kernel) {
      keep.push_back(p);
      continue;
    }
    HIP_OK(hipEventSynchronize(p.b));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, p.a, p.b));
    tot += ms;
    ++cnt;
    ctx->pool.push_back(p.a);
    ctx->pool.push_back(p.b);
  }
  ctx->pairs.swap(keep);
  *ms_total = tot;
  *launches = cnt;
  return GM_OK;
}

int gm_nmax(double size) { return (int)rint(2.0 + size + 4.0 * pow(size, 1.0 / 3.0)); }

int gm_coef_offsets(gm_ctx* ctx, int64_t n, const double* size, int64_t* off, gm_stream stream) {
  if (!ctx || n < 0 || (n > 0 && (!size || !off)))
    return gm_fail(GM_EINVAL, "gm_coef_offsets: bad arguments");
  HIP_OK(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  // counts into off[1..n], then an in-place inclusive scan; off[0] is cleared
  // by a kernel (no memset node when the caller captures the stream)
  hipLaunchKernelGGL(gm::k_zero_i64, dim3(1), dim3(64), 0, s, off, 1);
  if (n == 0) return GM_OK;
  hipLaunchKernelGGL(gm::k_nmax_of, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, size,
                     off + 1);
  HIP_OK(hipGetLastError());
  return gm::scan_inclusive_i64(ctx, off + 1, n, s);
}

int gm_mie_batch(gm_ctx* ctx, int64_t n, const double* x, const double* y, const double* eps,
                 const double* mu, const double* eps2, double size_max, const int64_t* coef_off,
                 double* an, double* bn, double* props, int32_t* nmax_out, uint32_t flags,
                 gm_stream stream) {
  int rc = mie_batch_check(ctx, n, x, y, eps, eps2, size_max, coef_off, an, bn, "gm_mie_batch");
  if (rc || n == 0) return rc;
  HIP_OK(hipSetDevice(ctx->device));
  const int nb = gm::kClasses * gm::kBuckets;
  if ((rc = ctx->reserve(ctx->order, sizeof(int64_t) * n))) return rc;
  if ((rc = ctx->reserve(ctx->key, sizeof(int32_t) * n))) return rc;
  if ((rc = ctx->reserve(ctx->hist, sizeof(int64_t) * (2 * nb + gm::kQueueSlot + 1)))) return rc;
  MieWs ws;
  ws.order = (int64_t*)ctx->order.ptr;
  ws.key = (int32_t*)ctx->key.ptr;
  ws.hist = (int64_t*)ctx->hist.ptr;
  ws.status = (int32_t*)ctx->status.ptr;
  const int rows_cap = std::max(1, gm_nmax(size_max));
  MieGrids g = mie_grids(ctx, n, rows_cap, ctx->scratch_budget);
  size_t need = g.homog.second;
  if (y) need = std::max(need, std::max(g.stable.second, g.literal.second));
  if ((rc = ctx->reserve(ctx->scratch, need))) return rc;
  ws.slots = (double*)ctx->scratch.ptr;
  return mie_batch_run(ctx, n, x, y, eps, mu, eps2, rows_cap, coef_off, an, bn, props, nmax_out,
                       flags, ws, g, (hipStream_t)stream);
}

int gm_mie_batch_noalloc(gm_ctx* ctx, int64_t n, const double* x, const double* y,
                         const double* eps, const double* mu, const double* eps2,
                         double size_max, const int64_t* coef_off, double* an, double* bn,
                         double* props, int32_t* nmax_out, uint32_t flags, int32_t* status,
                         gm_stream stream) {
  int rc = mie_batch_check(ctx, n, x, y, eps, eps2, size_max, coef_off, an, bn,
                           "gm_mie_batch_noalloc");
  if (rc) return rc;
  if (!status) return gm_fail(GM_EINVAL, "gm_mie_batch_noalloc: status is NULL");
  HIP_OK(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)noalloc_scratch(ctx->device);
  if (!base)
    return gm_fail(GM_ERUNTIME, "gm_mie_batch_noalloc: no static scratch on device %d",
                   ctx->device);
  // the call's flags start at zero (a fill kernel: no memset node if captured)
  hipLaunchKernelGGL(gm::k_zero_i32, dim3(1), dim3(64), 0, s, status, 1);
  HIP_OK(hipGetLastError());
  if (n == 0) return GM_OK;
  // layout of the static array: plan histogram/cursors/ranges, sort keys,
  // sorted order, then the recurrence slots in what is left
  const int nb = gm::kClasses * gm::kBuckets;
  const size_t hist_bytes = sizeof(int64_t) * (2 * nb + gm::kQueueSlot + 1);
  auto up = [](size_t v) { return (v + 255) / 256 * 256; };
  const size_t plan_bytes = up(hist_bytes) + up(sizeof(int32_t) * (size_t)n) +
                            up(sizeof(int64_t) * (size_t)n);
  const int rows_cap = std::max(1, gm_nmax(size_max));
  const int row_doubles = y ? gm::kCoatedRowDoubles : gm::kHomogRowDoubles;
  // one recurrence slot of the largest kind any of the kernels may take
  const size_t slot_bytes =
      std::max((size_t)rows_cap * 64 * row_doubles * sizeof(double),
               (size_t)gm::kSmallLcap * 64 * gm::kHomogRowDoubles * sizeof(double));
  if (plan_bytes + slot_bytes > kNoAllocBytes)
    return gm_fail(GM_EINVAL,
                   "gm_mie_batch_noalloc: %lld points with series length %d need %zu bytes of "
                   "scratch for the plan and one recurrence slot; the static scratch holds %zu",
                   (long long)n, rows_cap, plan_bytes + slot_bytes, kNoAllocBytes);
  MieWs ws;
  ws.hist = (int64_t*)base;
  ws.key = (int32_t*)(base + up(hist_bytes));
  ws.order = (int64_t*)(base + up(hist_bytes) + up(sizeof(int32_t) * (size_t)n));
  ws.status = status;
  ws.slots = (double*)(base + plan_bytes);
  MieGrids g = mie_grids(ctx, n, rows_cap, kNoAllocBytes - plan_bytes);
  return mie_batch_run(ctx, n, x, y, eps, mu, eps2, rows_cap, coef_off, an, bn, props, nmax_out,
                       flags, ws, g, s);
}

}  // extern "C"
