// MLP surrogate of the LUT (bulk optics from 6 inputs) on gfx950: forward,
// backward and Adam, hand-written.  No counterpart in the reference.
//
// Network: dims[0] -> dims[1] -> ... -> dims[L], SiLU after every layer but
// the last.  Parameters are one FP32 vector, per layer W [out][in] row-major
// then b [out].  The model's weights are BF16: every product uses bf16(W)
// (round to nearest even), while training keeps FP32 master copies.
// Activations stay FP32 and enter the matrix cores split into two BF16 terms
// (a = a_hi + a_lo), so a layer computes  a_hi W + a_lo W  with FP32
// accumulation (v_mfma_f32_16x16x32_bf16) -- the activation rounding is
// ~2^-17 instead of BF16's 2^-9.  Weight gradients use a three-term split of
// both operands.  A context may instead carry the weights as two BF16 terms
// (GM_TUNE_MLP_WEIGHT_TERMS = 2: W = W_hi + W_lo, three products per layer).
//
// GEMM kernel: C[m][n] = sum_k A(m, k) B(k, n), operands given by strides so
// one kernel serves the forward (X W^T), the activation gradient (dZ W) and
// the weight gradient (dZ^T A); tiles of 64 x 64 x 32 are staged through LDS
// (as BF16 hi/lo planes, K contiguous) whichever operand dimension is
// contiguous in memory.  Four wavefronts, each a 32 x 32 block of 2 x 2 MFMA
// tiles.  f32/bf16 16x16 MFMA results land as D[4g + r][j] in lane j + 16g.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "gm_ctx.h"

typedef __bf16 bf16x8_ __attribute__((ext_vector_type(8)));
typedef float float4_ __attribute__((ext_vector_type(4)));

namespace gm {

constexpr int kBM = 64, kBN = 64, kBK = 32;
constexpr int kLdk = kBK + 8;   // padded LDS row (80 bytes, keeps 16-byte alignment)
constexpr int kMlpMaxLayers = 16;
constexpr int kMlpMaxWidth = 4096;

enum { kEpiFwd = 0, kEpiBwdAct = 1, kEpiStore = 2 };

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// 8 consecutive floats at p (16-byte aligned) -> BF16 hi (and lo) planes, one
// 16-byte LDS store each; zeros when !ok.
template <int S>
__device__ __forceinline__ void stage8(const float* __restrict__ p, bool ok,
                                       unsigned short* __restrict__ dh,
                                       unsigned short* __restrict__ dl) {
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
  if (ok) {
    a = *(const float4*)p;
    b = *(const float4*)(p + 4);
  }
  const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  union { uint4 q; unsigned short s[8]; } h, l;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    h.s[e] = f2bf(v[e]);
    if (S == 2) l.s[e] = f2bf(v[e] - bf2f(h.s[e]));
  }
  *(uint4*)dh = h.q;
  if (S == 2) *(uint4*)dl = l.q;
}

struct GemmArgs {
  int M, N, K;
  const float* A;
  int64_t sam, sak;     // A(m, k) = A[m * sam + k * sak]
  const float* B;
  int64_t sbk, sbn;     // B(k, n) = B[k * sbk + n * sbn]
  float* C;
  int64_t ldc;          // C[m * ldc + n]
  const float* bias;    // kEpiFwd: + bias[n]
  float* Z;             // kEpiFwd: pre-activation store (may be null)
  int act;              // kEpiFwd: C = SiLU(z) (else z)
  const float* Zprev;   // kEpiBwdAct: C = acc * SiLU'(Zprev[m * ldc + n])
  int k_chunk;          // split-K: blockIdx.z covers [z * k_chunk, (z + 1) * k_chunk)
  int64_t split_stride; // kEpiStore: C offset of split z
};

// SA / SB: BF16 terms per operand (1 = plain BF16 rounding, 2 = hi + lo).
template <int SA, int SB, int EPI>
__global__ __launch_bounds__(256) void k_mlp_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) unsigned short sA[SA][kBM][kLdk];
  __shared__ __attribute__((aligned(16))) unsigned short sB[SB][kBN][kLdk];
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
  const int kbeg = blockIdx.z * g.k_chunk;
  const int kend = min(g.K, kbeg + g.k_chunk);
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  const int i = lane & 15, kq = lane >> 4;
  float4_ acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (float4_){0.f, 0.f, 0.f, 0.f};

  for (int k0 = kbeg; k0 < kend; k0 += kBK) {
    // ---- stage A (64 x 32) and B (32 x 64) as BF16 planes [row][k].  Where
    // the operand is K-contiguous, 16-byte aligned and the K-step is whole,
    // a thread moves 8 consecutive k of one row with two 16-byte loads and
    // one 16-byte LDS store per plane; otherwise element by element.
    const bool full_k = k0 + kBK <= kend;
    if (full_k && g.sak == 1 && (g.sam & 3) == 0 && (((uintptr_t)g.A) & 15) == 0) {
      const int mm = t >> 2, kc = (t & 3) * 8;
      stage8<SA>(g.A + (int64_t)(m0 + mm) * g.sam + k0 + kc, m0 + mm < g.M, &sA[0][mm][kc],
                 &sA[SA - 1][mm][kc]);
    } else {
      int r, c, dr, dc;   // element e at (r + e*dr, c + e*dc) of the tile
      if (g.sak == 1) { r = t >> 2; c = (t & 3) * 8; dr = 0; dc = 1; }
      else            { c = t >> 3; r = (t & 7) * 8; dr = 1; dc = 0; }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int mm = r + e * dr, kk = c + e * dc;
        const int gm_ = m0 + mm, gk = k0 + kk;
        const float v = (gm_ < g.M && gk < kend) ? g.A[(int64_t)gm_ * g.sam + (int64_t)gk * g.sak] : 0.f;
        const unsigned short h = f2bf(v);
        sA[0][mm][kk] = h;
        if (SA == 2) sA[SA - 1][mm][kk] = f2bf(v - bf2f(h));
      }
    }
    if (full_k && g.sbk == 1 && (g.sbn & 3) == 0 && (((uintptr_t)g.B) & 15) == 0) {
      const int nn = t >> 2, kc = (t & 3) * 8;
      stage8<SB>(g.B + (int64_t)(n0 + nn) * g.sbn + k0 + kc, n0 + nn < g.N, &sB[0][nn][kc],
                 &sB[SB - 1][nn][kc]);
    } else {
      int r, c, dr, dc;   // r: column n, c: k
      if (g.sbk == 1) { r = t >> 2; c = (t & 3) * 8; dr = 0; dc = 1; }
      else            { c = t >> 3; r = (t & 7) * 8; dr = 1; dc = 0; }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int nn = r + e * dr, kk = c + e * dc;
        const int gn = n0 + nn, gk = k0 + kk;
        const float v = (gn < g.N && gk < kend) ? g.B[(int64_t)gk * g.sbk + (int64_t)gn * g.sbn] : 0.f;
        const unsigned short h = f2bf(v);
        sB[0][nn][kk] = h;
        if (SB == 2) sB[SB - 1][nn][kk] = f2bf(v - bf2f(h));
      }
    }
    __syncthreads();
    // ---- 2 x 2 MFMA tiles per wavefront
    union U { uint4 q; bf16x8_ v; };
    U ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      ah[a].q = *(const uint4*)&sA[0][wm + 16 * a + i][8 * kq];
      if (SA == 2) al[a].q = *(const uint4*)&sA[SA - 1][wm + 16 * a + i][8 * kq];
      bh[a].q = *(const uint4*)&sB[0][wn + 16 * a + i][8 * kq];
      if (SB == 2) bl[a].q = *(const uint4*)&sB[SB - 1][wn + 16 * a + i][8 * kq];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a].v, bh[b].v, acc[a][b], 0, 0, 0);
        if (SA == 2)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a].v, bh[b].v, acc[a][b], 0, 0, 0);
        if (SB == 2)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a].v, bl[b].v, acc[a][b], 0, 0, 0);
      }
    __syncthreads();
  }
  // ---- epilogue
  float* C = g.C + (EPI == kEpiStore ? (int64_t)blockIdx.z * g.split_stride : 0);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = n0 + wn + 16 * b + i;
      if (n >= g.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * a + 4 * kq + r;
        if (m >= g.M) continue;
        const int64_t o = (int64_t)m * g.ldc + n;
        const float v = acc[a][b][r];
        if (EPI == kEpiFwd) {
          const float z = v + g.bias[n];
          if (g.Z) g.Z[o] = z;
          C[o] = g.act ? silu(z) : z;
        } else if (EPI == kEpiBwdAct) {
          C[o] = v * dsilu(g.Zprev[o]);
        } else {
          C[o] = v;
        }
      }
    }
}

// out[i] = sum_q part[q * n + i], q in order.
__global__ void k_mlp_reduce(int n_split, int64_t n, const float* __restrict__ part,
                             float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int q = 0; q < n_split; ++q) s += part[(int64_t)q * n + i];
  out[i] = s;
}

// Column sums of X [rows][cols] by row chunks: part[chunk][col].
__global__ __launch_bounds__(256) void k_mlp_colsum(int64_t rows, int cols, int64_t chunk,
                                                    const float* __restrict__ X,
                                                    float* __restrict__ part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int sub = threadIdx.x >> 6;
  const int64_t r0 = (int64_t)blockIdx.y * chunk;
  const int64_t r1 = min(rows, r0 + chunk);
  float s = 0.f;
  if (c < cols)
    for (int64_t r = r0 + sub; r < r1; r += 4) s += X[r * cols + c];
  red[sub][threadIdx.x & 63] = s;
  __syncthreads();
  if (sub == 0 && c < cols)
    part[(int64_t)blockIdx.y * cols + c] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// dY = 2 cw[j] (Y - T) / n;  lpart[block] = sum cw[j] (Y - T)^2 over the block.
__global__ __launch_bounds__(256) void k_mlp_mse(int64_t n, int n_out, const float* __restrict__ Y,
                                                 const float* __restrict__ T,
                                                 const float* __restrict__ cw,
                                                 float* __restrict__ dY,
                                                 float* __restrict__ lpart) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < n * n_out) {
    const int j = (int)(i % n_out);
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * cw[j] * d / (float)n;
    l = cw[j] * d * d;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

__global__ void k_mlp_loss_final(int n_part, int64_t n, const float* __restrict__ lpart,
                                 float* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int q = 0; q < n_part; ++q) s += lpart[q];
  loss[0] = (float)(s / (double)n);
}

__global__ void k_mlp_adam(int64_t n, float* __restrict__ p, const float* __restrict__ grad,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

}  // namespace gm

namespace {

struct Net {
  int L = 0;
  int dims[gm::kMlpMaxLayers + 1];
  int64_t woff[gm::kMlpMaxLayers], boff[gm::kMlpMaxLayers];
  int64_t n_params = 0;
  int max_width = 0;
};

int make_net(int n_layers, const int* dims, Net& net, const char* who) {
  if (n_layers < 1 || n_layers > gm::kMlpMaxLayers || !dims)
    return gm_fail(GM_EINVAL, "%s: n_layers must be in [1, %d]", who, gm::kMlpMaxLayers);
  net.L = n_layers;
  int64_t off = 0;
  net.max_width = 0;
  for (int l = 0; l <= n_layers; ++l) {
    if (dims[l] < 1 || dims[l] > gm::kMlpMaxWidth)
      return gm_fail(GM_EINVAL, "%s: layer width %d out of [1, %d]", who, dims[l], gm::kMlpMaxWidth);
    net.dims[l] = dims[l];
    net.max_width = std::max(net.max_width, dims[l]);
  }
  for (int l = 0; l < n_layers; ++l) {
    net.woff[l] = off;
    off += (int64_t)dims[l] * dims[l + 1];
    net.boff[l] = off;
    off += dims[l + 1];
  }
  net.n_params = off;
  return GM_OK;
}

template <int SA, int SB, int EPI>
int gemm(const gm::GemmArgs& a, int n_split, hipStream_t s) {
  dim3 grid((unsigned)((a.M + gm::kBM - 1) / gm::kBM), (unsigned)((a.N + gm::kBN - 1) / gm::kBN),
            (unsigned)n_split);
  hipLaunchKernelGGL((gm::k_mlp_gemm<SA, SB, EPI>), grid, dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "k_mlp_gemm: %s", hipGetErrorString(e));
  return GM_OK;
}

gm::GemmArgs args0() {
  gm::GemmArgs a;
  a.M = a.N = a.K = 0;
  a.A = a.B = nullptr;
  a.sam = a.sak = a.sbk = a.sbn = 0;
  a.C = nullptr;
  a.ldc = 0;
  a.bias = nullptr;
  a.Z = nullptr;
  a.act = 0;
  a.Zprev = nullptr;
  a.k_chunk = 1 << 30;
  a.split_stride = 0;
  return a;
}

// Forward of layer l: C[n][out] from Ain[n][in]; Z (optional) keeps z.
int layer_forward(const Net& net, int l, int64_t n, const float* params, const float* Ain,
                  float* C, float* Z, int wterms, hipStream_t s) {
  gm::GemmArgs a = args0();
  const int in = net.dims[l], out = net.dims[l + 1];
  a.M = (int)n;
  a.N = out;
  a.K = in;
  a.A = Ain;
  a.sam = in;
  a.sak = 1;
  a.B = params + net.woff[l];   // B(k, j) = W[j][k]
  a.sbk = 1;
  a.sbn = in;
  a.C = C;
  a.ldc = out;
  a.bias = params + net.boff[l];
  a.Z = Z;
  a.act = l < net.L - 1;
  return wterms == 2 ? gemm<2, 2, gm::kEpiFwd>(a, 1, s) : gemm<2, 1, gm::kEpiFwd>(a, 1, s);
}

}  // namespace

extern "C" {

// See include/geosmie.h.
int64_t gm_mlp_param_count(int n_layers, const int* dims) {
  Net net;
  if (make_net(n_layers, dims, net, "gm_mlp_param_count")) return -1;
  return net.n_params;
}

int gm_mlp_forward(gm_ctx* ctx, int n_layers, const int* dims, int64_t n, const float* params,
                   const float* x, float* y, gm_stream stream) {
  Net net;
  int rc = make_net(n_layers, dims, net, "gm_mlp_forward");
  if (rc) return rc;
  if (!ctx || n < 0) return gm_fail(GM_EINVAL, "gm_mlp_forward: bad arguments");
  if (n == 0) return GM_OK;
  if (n > 0x7fffffffLL / 2) return gm_fail(GM_EINVAL, "gm_mlp_forward: batch too large");
  if (!params || !x || !y) return gm_fail(GM_EINVAL, "gm_mlp_forward: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  const size_t act = (size_t)n * net.max_width;
  if ((rc = ctx->reserve(ctx->mlp_ws, 2 * act * sizeof(float)))) return rc;
  float* buf[2] = {(float*)ctx->mlp_ws.ptr, (float*)ctx->mlp_ws.ptr + act};
  size_t slot;
  ctx->mark_begin(gm::kKMlp, s, &slot);
  const float* in = x;
  for (int l = 0; l < net.L; ++l) {
    float* out = l == net.L - 1 ? y : buf[l & 1];
    if ((rc = layer_forward(net, l, n, params, in, out, nullptr, ctx->tune_mlp_wterms, s))) return rc;
    in = out;
  }
  ctx->mark_end(slot, s);
  return GM_OK;
}

int gm_mlp_grad(gm_ctx* ctx, int n_layers, const int* dims, int64_t n, const float* params,
                const float* x, const float* target, const float* out_weight, float* grads,
                float* loss, gm_stream stream) {
  Net net;
  int rc = make_net(n_layers, dims, net, "gm_mlp_grad");
  if (rc) return rc;
  if (!ctx || n < 1) return gm_fail(GM_EINVAL, "gm_mlp_grad: bad arguments");
  if (n > 0x7fffffffLL / 2) return gm_fail(GM_EINVAL, "gm_mlp_grad: batch too large");
  if (!params || !x || !target || !out_weight || !grads || !loss)
    return gm_fail(GM_EINVAL, "gm_mlp_grad: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  const int L = net.L;
  // workspace: A_l and Z_l for every layer, two gradient buffers, split-K
  // partials and the loss partials
  std::vector<size_t> aoff(L), zoff(L);
  size_t off = 0;
  for (int l = 0; l < L; ++l) {
    aoff[l] = off;
    off += (size_t)n * net.dims[l + 1];
    zoff[l] = off;
    off += (size_t)n * net.dims[l + 1];
  }
  const size_t g0 = off, g1 = g0 + (size_t)n * net.max_width;
  off = g1 + (size_t)n * net.max_width;
  // split-K over the batch for the weight gradients: about 512 blocks per layer
  int n_split = 1;
  int64_t k_chunk = n;
  {
    int64_t want = std::max<int64_t>(1, std::min<int64_t>(64, n / 256));
    k_chunk = ((n + want - 1) / want + gm::kBK - 1) / gm::kBK * gm::kBK;
    n_split = (int)((n + k_chunk - 1) / k_chunk);
  }
  const size_t part = off;
  size_t max_w = 0;
  for (int l = 0; l < L; ++l) max_w = std::max(max_w, (size_t)net.dims[l] * net.dims[l + 1]);
  const int64_t cs_chunk = 1024;
  const int64_t cs_n = (n + cs_chunk - 1) / cs_chunk;
  off += std::max((size_t)n_split * max_w, (size_t)cs_n * net.max_width);
  const size_t lp = off;
  const int64_t n_out_el = n * net.dims[L];
  const int64_t mse_blocks = (n_out_el + 255) / 256;
  off += mse_blocks;
  if ((rc = ctx->reserve(ctx->mlp_ws, off * sizeof(float)))) return rc;
  float* W = (float*)ctx->mlp_ws.ptr;
  size_t slot;
  ctx->mark_begin(gm::kKMlp, s, &slot);
  // forward
  const float* in = x;
  for (int l = 0; l < L; ++l) {
    if ((rc = layer_forward(net, l, n, params, in, W + aoff[l], W + zoff[l], ctx->tune_mlp_wterms, s)))
      return rc;
    in = W + aoff[l];
  }
  // loss and dY
  float* gcur = W + g0;
  float* gnext = W + g1;
  hipLaunchKernelGGL(gm::k_mlp_mse, dim3((unsigned)mse_blocks), dim3(256), 0, s, n, net.dims[L],
                     (const float*)(W + aoff[L - 1]), target, out_weight, gcur, W + lp);
  hipLaunchKernelGGL(gm::k_mlp_loss_final, dim3(1), dim3(64), 0, s, (int)mse_blocks, n,
                     (const float*)(W + lp), loss);
  // backward
  for (int l = L - 1; l >= 0; --l) {
    const int fin = net.dims[l], fout = net.dims[l + 1];
    const float* Aprev = l == 0 ? x : W + aoff[l - 1];
    // dW[o][i] = sum_b dZ[b][o] Aprev[b][i]
    gm::GemmArgs a = args0();
    a.M = fout;
    a.N = fin;
    a.K = (int)n;
    a.A = gcur;
    a.sam = 1;
    a.sak = fout;
    a.B = Aprev;
    a.sbk = fin;
    a.sbn = 1;
    a.C = W + part;
    a.ldc = fin;
    a.k_chunk = (int)k_chunk;
    a.split_stride = (int64_t)fout * fin;
    if ((rc = gemm<2, 2, gm::kEpiStore>(a, n_split, s))) return rc;
    const int64_t nw = (int64_t)fout * fin;
    hipLaunchKernelGGL(gm::k_mlp_reduce, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s,
                       n_split, nw, (const float*)(W + part), grads + net.woff[l]);
    // db[o] = sum_b dZ[b][o]
    hipLaunchKernelGGL(gm::k_mlp_colsum, dim3((unsigned)((fout + 63) / 64), (unsigned)cs_n),
                       dim3(256), 0, s, n, fout, cs_chunk, (const float*)gcur, W + part);
    hipLaunchKernelGGL(gm::k_mlp_reduce, dim3((unsigned)((fout + 255) / 256)), dim3(256), 0, s,
                       (int)cs_n, (int64_t)fout, (const float*)(W + part), grads + net.boff[l]);
    if (l > 0) {
      // dZ_{l-1}[b][i] = (sum_o dZ[b][o] W[o][i]) * SiLU'(Z_{l-1}[b][i])
      gm::GemmArgs d = args0();
      d.M = (int)n;
      d.N = fin;
      d.K = fout;
      d.A = gcur;
      d.sam = fout;
      d.sak = 1;
      d.B = params + net.woff[l];
      d.sbk = fin;
      d.sbn = 1;
      d.C = gnext;
      d.ldc = fin;
      d.Zprev = W + zoff[l - 1];
      rc = ctx->tune_mlp_wterms == 2 ? gemm<2, 2, gm::kEpiBwdAct>(d, 1, s)
                                     : gemm<2, 1, gm::kEpiBwdAct>(d, 1, s);
      if (rc) return rc;
      std::swap(gcur, gnext);
    }
  }
  ctx->mark_end(slot, s);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "gm_mlp_grad: %s", hipGetErrorString(e));
  return GM_OK;
}

int gm_mlp_adam(gm_ctx* ctx, int64_t n_params, float* params, const float* grads, float* m,
                float* v, float lr, float beta1, float beta2, float eps, int64_t step,
                gm_stream stream) {
  if (!ctx || n_params < 0 || step < 1) return gm_fail(GM_EINVAL, "gm_mlp_adam: bad arguments");
  if (n_params == 0) return GM_OK;
  if (!params || !grads || !m || !v) return gm_fail(GM_EINVAL, "gm_mlp_adam: NULL array");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, (double)step)));
  const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, (double)step)));
  hipLaunchKernelGGL(gm::k_mlp_adam, dim3((unsigned)((n_params + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n_params, params, grads, m, v, lr, beta1, beta2, eps,
                     c1, c2);
  e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(GM_ERUNTIME, "k_mlp_adam: %s", hipGetErrorString(e));
  return GM_OK;
}

}  // extern "C"
