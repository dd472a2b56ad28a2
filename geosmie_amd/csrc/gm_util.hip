// Small device utilities: int64 prefix scan, n_max per point.
#include <hip/hip_runtime.h>

#include "gm_ctx.h"

namespace gm {

constexpr int kScanBlock = 1024;

// Per-block inclusive scan in LDS; block totals to sums[blockIdx].
__global__ __launch_bounds__(kScanBlock) void k_scan_block(int64_t* __restrict__ d, int64_t n,
                                                           int64_t* __restrict__ sums) {
  __shared__ int64_t buf[kScanBlock];
  int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  buf[threadIdx.x] = i < n ? d[i] : 0;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int64_t v = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
    __syncthreads();
    buf[threadIdx.x] += v;
    __syncthreads();
  }
  if (i < n) d[i] = buf[threadIdx.x];
  if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = buf[threadIdx.x];
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(int64_t* __restrict__ d, int64_t n,
                                                         const int64_t* __restrict__ sums) {
  int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  if (blockIdx.x > 0 && i < n) d[i] += sums[blockIdx.x - 1];
}

int scan_inclusive_i64(gm_ctx* ctx, int64_t* data, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  int64_t nblk = (n + kScanBlock - 1) / kScanBlock;
  // scratch for this level's block sums, then recurse on them
  std::vector<int64_t> levels;
  size_t need = 0;
  for (int64_t m = nblk; ; m = (m + kScanBlock - 1) / kScanBlock) {
    need += (size_t)m;
    if (m <= 1) break;
  }
  int rc = ctx->reserve(ctx->scan_tmp, need * sizeof(int64_t) + 64);
  if (rc) return rc;
  int64_t* sums = (int64_t*)ctx->scan_tmp.ptr;
  // iterative: scan data -> sums0; scan sums0 -> sums1; ...; then add back down
  std::vector<int64_t*> ptrs;
  std::vector<int64_t> lens;
  int64_t* cur = data;
  int64_t len = n;
  int64_t* next = sums;
  while (true) {
    int64_t b = (len + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)b), dim3(kScanBlock), 0, s, cur, len, next);
    ptrs.push_back(cur);
    lens.push_back(len);
    if (b <= 1) break;
    cur = next;
    len = b;
    next = next + b;
  }
  for (int lv = (int)ptrs.size() - 2; lv >= 0; --lv) {
    int64_t b = (lens[lv] + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)b), dim3(kScanBlock), 0, s, ptrs[lv], lens[lv],
                       ptrs[lv + 1]);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return gm_fail(1, "scan: %s", hipGetErrorString(e));
  return 0;
}

__global__ void k_nmax_i32(int64_t n, const double* __restrict__ size, int32_t* __restrict__ out) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) out[p] = (int)rint(2.0 + size[p] + 4.0 * pow(size[p], 1.0 / 3.0));
}

}  // namespace gm
