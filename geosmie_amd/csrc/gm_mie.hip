// Mie coefficient kernels for gfx950: homogeneous and coated spheres, FP64.
//
// One (size, m) point per lane.  Points are first classified (homogeneous,
// coated with a real or weakly absorbing shell, coated with an absorbing
// shell; the homogeneous/coated split follows the dispatch of
// src/pymiecoated/pymiecoated/mie_coeffs.py:53-73) and counting-sorted by
// recurrence length, longest first, so that the 64 lanes of a wavefront run
// nearly equal trip counts and the hardware dispatcher hands out the long
// wavefronts first.  The coefficient kernels are persistent: each wavefront
// owns one scratch slot (reused for every chunk it processes, so the slot
// stays L2/MALL resident) and walks the sorted list with a grid stride.
//
// Scratch layout per slot: plane[q][row n][lane] so every row access of a
// wavefront is one coalesced 512 B (double) or 1 KiB (complex) transaction.
//
// Series lengths follow mie_coeffs.py:99,101 (homogeneous) and :203-205
// (coated): nmax = round(2 + s + 4 s^(1/3)), nmx = round(max(nmax,|z|)+16)
// with Python's half-to-even rounding (rint).  The log-derivative D_n is the
// reference's downward recurrence from D_nmx = 0 (mie_coeffs.py:118-121),
// reproduced exactly, including its start index.
//
// Riccati-Bessel functions psi_n = s j_n(s), chi_n = -s y_n(s) (which the
// reference takes from scipy jv/yv, mie_coeffs.py:105-113, 220-227) are
// computed here: chi_n by upward recurrence (dominant, stable); psi_n by
// upward recurrence while n <= |s| (oscillatory region) and, above it, from
// psi_n = psi_{n-1} / (D_n(s) + n/s) with D_n(s) from a separate downward
// recurrence started 16 + 8 ceil(|s|^(1/3)) above nmax (converged), which is
// stable where psi_n decays.
#include "gm_complex.h"
#include "gm_internal.h"

namespace gm {

// ---------------------------------------------------------------------------
// Series-length rules
// ---------------------------------------------------------------------------
__device__ __forceinline__ int nmax_rule(double s) {
  return (int)rint(2.0 + s + 4.0 * pow(s, 1.0 / 3.0));
}
__device__ __forceinline__ int nmx_rule(int nmax, double zabs) {
  return (int)rint(fmax((double)nmax, zabs) + 16.0);
}
// Switch row of a psi chain with no ratio rows (|z| >= n_max): above every
// row, so "row k uses the ratio chain" is k > nsw alone.
constexpr int kNoRatio = 1 << 30;
// Start of the downward psi-ratio chains: rows n <= nmax agree to < 1e-15
// with any later start from nmax + 8 + 4 ceil(|z|^(1/3)) on (checked over
// |z| = 0.01..2000 and the refractive indices of the workloads); the margin
// below is half again that.
__device__ __forceinline__ int ratio_start(int nmax, double zabs) {
  return nmax + 12 + 6 * (int)ceil(cbrt(zabs));
}
// Whether a downward D_n recurrence started at nmx has converged (relative
// error < ~1e-17) by the time it reaches n <= |z|: the error decays like
// exp(-(4/3) zeta^(3/2)) across the Airy region, zeta = (nmx - |z|) / (|z|/2)^(1/3).
__device__ __forceinline__ bool converged_start(int nmx, double zabs) {
  return (double)nmx - zabs >= 8.0 * cbrt(zabs);
}

// Effective per-point parameters after the reference's dispatch.
struct PointDesc {
  int cls;       // 0 homogeneous, 1 coated (|Im m2| y <= 1), 2 coated (absorbing shell)
  double size;   // x for homogeneous (after dispatch), y for coated
  double xcore;  // coated core size
  cd eps, mu, eps2;
};

__device__ __forceinline__ PointDesc describe(int64_t p, const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ eps,
                                              const double* __restrict__ mu,
                                              const double* __restrict__ eps2) {
  PointDesc d;
  d.eps = cd{eps[2 * p], eps[2 * p + 1]};
  d.mu = mu ? cd{mu[2 * p], mu[2 * p + 1]} : cd{1.0, 0.0};
  double xx = x[p];
  if (!y) {
    d.cls = 0;
    d.size = xx;
    d.xcore = xx;
    d.eps2 = d.eps;
    return d;
  }
  double yy = y[p];
  cd e2 = cd{eps2[2 * p], eps2[2 * p + 1]};
  d.eps2 = e2;
  d.size = yy;
  d.xcore = xx;
  // mie_coeffs.py:66-71
  if (xx == yy || (d.eps.re == e2.re && d.eps.im == e2.im)) {
    d.cls = 0;
  } else if (xx == 0.0) {
    d.cls = 0;
    d.eps = e2;
  } else {
    // The cancellation-free shell combination (k_mie_coated_stable) suits real
    // and weakly absorbing shells; strongly absorbing shells (no real zeros of
    // psi_n(m2 y), but e^{Im}-sized psi/chi products) use the reference's
    // ratio form (k_mie_coated_literal).
    cd m2 = csqrt_(e2);
    d.cls = (fabs(m2.im) * yy <= 1.0) ? 1 : 2;
  }
  return d;
}

// Work estimate used as the sort key (number of complex recurrence steps):
// the n_max and nmx rules on the moduli alone (|sqrt(e)| = sqrt|e|), with
// cbrt for pow(s, 1/3) -- no complex root or FP64 pow; only the bucket
// (cost >> kBucketShift) matters.
__device__ __forceinline__ int point_cost(const PointDesc& d) {
  const double s = d.size;
  const double nmax = rint(2.0 + s + 4.0 * cbrt(s));
  double zabs;
  if (d.cls == 0) {
    const cd em = d.eps * d.mu;
    zabs = sqrt(hypot(em.re, em.im)) * s;
  } else {
    zabs = fmax(sqrt(hypot(d.eps.re, d.eps.im)), sqrt(hypot(d.eps2.re, d.eps2.im))) * s;
  }
  const double nmx = rint(fmax(nmax, zabs) + 16.0);
  const double c = d.cls == 0 ? nmx + 2.0 * nmax
                              : (d.cls == 1 ? nmx + 4.0 * nmax : 3.0 * nmx + 6.0 * nmax);
  return (int)fmin(c, 1.0e9);
}

// ---------------------------------------------------------------------------
// Plan: classify, n_max, counting sort by cost (descending) per class
// ---------------------------------------------------------------------------
__global__ void k_zero_i64(int64_t* __restrict__ p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_zero_i32(int32_t* __restrict__ p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Block-local LDS histogram, then one global atomic per non-empty bucket.
// Each block covers per * 256 points (per <= kPlanPer, chosen from n: large
// batches amortise the histogram's zeroing and flush, small ones keep enough
// blocks in flight).  status bit 2: a coated point with mu != 1 (mie_coeffs.py:57-59).
__global__ __launch_bounds__(256) void k_plan_keys(
    int64_t n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ eps, const double* __restrict__ mu,
    const double* __restrict__ eps2, int32_t* __restrict__ key, int32_t* __restrict__ nmax_out,
    int64_t* __restrict__ hist, int32_t* __restrict__ status, int per) {
  __shared__ int32_t lh[kClasses * kBuckets];
  for (int i = threadIdx.x; i < kClasses * kBuckets; i += blockDim.x) lh[i] = 0;
  __syncthreads();
  const int64_t p0 = (int64_t)blockIdx.x * (256 * (int64_t)per);
  bool bad_mu = false;
  for (int q = 0; q < per; ++q) {
    const int64_t p = p0 + q * 256 + threadIdx.x;
    if (p >= n) break;
    PointDesc d = describe(p, x, y, eps, mu, eps2);
    int b = min(max(point_cost(d), 0) >> kBucketShift, kBuckets - 1);
    int k = d.cls * kBuckets + (kBuckets - 1 - b);  // descending cost inside a class
    key[p] = k;
    if (nmax_out) nmax_out[p] = nmax_rule(d.size);
    atomicAdd(&lh[k], 1);
    bad_mu |= (y && mu && (d.mu.re != 1.0 || d.mu.im != 0.0));
  }
  if (bad_mu) atomicOr(status, 2);
  __syncthreads();
  for (int i = threadIdx.x; i < kClasses * kBuckets; i += blockDim.x)
    if (lh[i]) atomicAdd((unsigned long long*)&hist[i], (unsigned long long)lh[i]);
}

// Exclusive scan of the kClasses*kBuckets histogram (one block of 256).
// cursor[k] = start of bucket k; range[2c], range[2c+1] = [begin, end) of class c.
// Each thread sums its run of buckets, the 256 run sums are scanned in LDS
// (Hillis-Steele, 8 steps), then each thread writes its run's cursors and the
// class boundaries that fall in it.
// Also splits the homogeneous class at the first bucket of cost below
// small_b (buckets descend along the order): ab = {begin, split, split, end},
// the points run with many lanes per point and those run with few.
__global__ __launch_bounds__(256) void k_plan_scan(const int64_t* __restrict__ hist,
                                                   int64_t* __restrict__ cursor,
                                                   int64_t* __restrict__ range, int small_b,
                                                   int64_t* __restrict__ ab) {
  __shared__ int64_t part[256];
  constexpr int per = (kClasses * kBuckets) / 256;
  static_assert(per * 256 == kClasses * kBuckets, "histogram must split into 256 runs");
  const int t = threadIdx.x;
  int64_t h[per];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < per; ++i) {
    h[i] = hist[t * per + i];
    s += h[i];
  }
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int64_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t acc = t == 0 ? 0 : part[t - 1];
#pragma unroll
  for (int i = 0; i < per; ++i) {
    const int j = t * per + i;
    cursor[j] = acc;
    if (j % kBuckets == 0) {
      const int c = j / kBuckets;
      range[2 * c] = acc;
      if (c > 0) range[2 * c - 1] = acc;
    }
    acc += h[i];
  }
  if (t == 255) range[2 * kClasses - 1] = acc;
  __syncthreads();
  if (t == 0 && ab) {
    const int key = kBuckets - min(max(small_b, 0), kBuckets);   // class 0's first small key
    const int64_t split = key < kBuckets ? cursor[key] : range[1];
    ab[0] = range[0];
    ab[1] = split;
    ab[2] = split;
    ab[3] = range[1];
    ab[4] = 0;   // the coated class's work queue (kQueueSlot)
  }
}

// Scatter with block-local ranks: one global atomic per (block, bucket),
// per * 256 points per block (per <= kPlanPer).
__global__ __launch_bounds__(256) void k_plan_scatter(int64_t n, const int32_t* __restrict__ key,
                                                      int64_t* __restrict__ cursor,
                                                      int64_t* __restrict__ order, int per) {
  __shared__ int32_t lcnt[kClasses * kBuckets];
  __shared__ int64_t lbase[kClasses * kBuckets];
  for (int i = threadIdx.x; i < kClasses * kBuckets; i += blockDim.x) lcnt[i] = 0;
  __syncthreads();
  const int64_t p0 = (int64_t)blockIdx.x * (256 * (int64_t)per);
  int kk[kPlanPer], rank[kPlanPer];
#pragma unroll
  for (int q = 0; q < kPlanPer; ++q) {
    const int64_t p = p0 + q * 256 + threadIdx.x;
    kk[q] = -1;
    rank[q] = 0;
    if (q < per && p < n) {
      kk[q] = key[p];
      rank[q] = atomicAdd(&lcnt[kk[q]], 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kClasses * kBuckets; i += blockDim.x)
    if (lcnt[i])
      lbase[i] = (int64_t)atomicAdd((unsigned long long*)&cursor[i], (unsigned long long)lcnt[i]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPlanPer; ++q)
    if (kk[q] >= 0) order[lbase[kk[q]] + rank[q]] = p0 + q * 256 + threadIdx.x;
}

// Identity order (no sort): every class kernel walks all points and keeps
// those of its own class.
__global__ void k_plan_identity(int64_t n, int64_t* __restrict__ order,
                                int64_t* __restrict__ range, int64_t* __restrict__ ab) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) order[p] = p;
  if (p == 0) {
    for (int c = 0; c < kClasses; ++c) {
      range[2 * c] = 0;
      range[2 * c + 1] = n;
    }
    if (ab) {   // unsorted: every point with many lanes per point
      ab[0] = 0;
      ab[1] = n;
      ab[2] = n;
      ab[3] = n;
      ab[4] = 0;   // the coated class's work queue (kQueueSlot)
    }
  }
}

// ---------------------------------------------------------------------------
// Efficiency accumulator, mie_props.py:28-70 (sequential FP64 sums over n)
// ---------------------------------------------------------------------------
struct Props {
  double sext, ssca, asy1, asy2;
  cd qb;
  cd ap, bp;  // a_{n-1}, b_{n-1}
  __device__ __forceinline__ void init() {
    sext = ssca = asy1 = asy2 = 0.0;
    qb = cd{0.0, 0.0};
    ap = bp = cd{0.0, 0.0};
  }
  // Orders visited one after another from n = 1: the asymmetry weights
  // c1(n) = (n-1)(n+1)/n = n - 1/n and c2(n) = (2n+1)/(n(n+1)) = 1/n + 1/(n+1)
  // from one reciprocal per order, 1/(n+1) carried into the next (rn holds 1/n;
  // init_seq sets it for n = 1).
  double rn;
  __device__ __forceinline__ void init_seq() {
    init();
    rn = 1.0;
  }
  __device__ __forceinline__ void add_seq(int n, cd a, cd b) {
    const double dn = (double)n;
    const double cn = 2.0 * dn + 1.0;
    const double rn1 = rcp_nr(dn + 1.0);
    sext = fma(cn, a.re + b.re, sext);
    ssca = fma(cn, norm2(a) + norm2(b), ssca);
    const double sg = (n & 1) ? -cn : cn;
    qb = qb + (a - b) * sg;
    // at n = 1, c1 = 0 and a_0 = b_0 = 0 leave asy1 unchanged (no branch)
    asy1 = fma(dn - rn, ap.re * a.re + ap.im * a.im + bp.re * b.re + bp.im * b.im, asy1);
    asy2 = fma(rn + rn1, a.re * b.re + a.im * b.im, asy2);
    rn = rn1;
    ap = a;
    bp = b;
  }
  __device__ __forceinline__ void add(int n, cd a, cd b) {
    double dn = (double)n;
    double cn = 2.0 * dn + 1.0;
    sext = fma(cn, a.re + b.re, sext);
    ssca = fma(cn, norm2(a) + norm2(b), ssca);
    double sg = (n & 1) ? -cn : cn;
    qb = qb + (a - b) * sg;
    {
      // at n = 1, c1 = 0 and a_0 = b_0 = 0 leave asy1 unchanged (no branch)
      double dm = dn - 1.0;  // previous index
      double c1 = dm * (dm + 2.0) * rcp_nr(dm + 1.0);
      asy1 = fma(c1, ap.re * a.re + ap.im * a.im + bp.re * b.re + bp.im * b.im, asy1);
    }
    double c2 = cn * rcp_nr(dn * (dn + 1.0));
    asy2 = fma(c2, a.re * b.re + a.im * b.im, asy2);
    ap = a;
    bp = b;
  }
  __device__ __forceinline__ void finish(double size, int64_t p, int64_t n_all,
                                         double* __restrict__ out) {
    double y2 = size * size;
    double qext = 2.0 * sext / y2;
    double qsca = 2.0 * ssca / y2;
    double qbv = (qb.re * qb.re + qb.im * qb.im) / y2;
    double asy = 4.0 / y2 * (asy1 + asy2) / qsca;
    out[0 * n_all + p] = qext;
    out[1 * n_all + p] = qsca;
    out[2 * n_all + p] = qext - qsca;
    out[3 * n_all + p] = qbv;
    out[4 * n_all + p] = asy;
    out[5 * n_all + p] = qbv / qsca;
  }
};

// ---------------------------------------------------------------------------
// Homogeneous sphere, mie_coeffs.py:83-130 / 132-180 + mie_props.py:28-70
// ---------------------------------------------------------------------------
// MINW: resident wavefronts per SIMD the register allocation must allow
// (1 = unconstrained; 3 and 4 trade a few spills for latency hiding).
template <int MINW>
__global__ __launch_bounds__(64, MINW) void k_mie_homog(
    const int64_t* __restrict__ order, const int64_t* __restrict__ range, int64_t n_all,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ eps,
    const double* __restrict__ mu, const double* __restrict__ eps2,
    const int64_t* __restrict__ coef_off, double* __restrict__ an, double* __restrict__ bn,
    double* __restrict__ props, int32_t* __restrict__ status, double* __restrict__ scratch,
    int rows_cap, int dn_extra) {
  const int lane = threadIdx.x;
  const int64_t begin = range[0], end = range[1];
  // slot planes: D (complex) and psi ratio (real)
  double* slot = scratch + (size_t)blockIdx.x * (size_t)rows_cap * 64 * 3;
  cd* Dp = reinterpret_cast<cd*>(slot);             // [row][lane]
  double* Rp = slot + (size_t)rows_cap * 64 * 2;    // [row][lane]

  for (int64_t base = begin + (int64_t)blockIdx.x * 64; base < end;
       base += (int64_t)gridDim.x * 64) {
    int64_t idx = base + lane;
    if (idx >= end) continue;
    int64_t p = order[idx];
    PointDesc d = describe(p, x, y, eps, mu, eps2);
    if (d.cls != 0) continue;
    const double s = d.size;
    const int nmax = nmax_rule(s);
    if (nmax > rows_cap) {
      atomicOr(status, 1);
      continue;
    }
    cd z = csqrt_(d.eps * d.mu) * s;
    cd m = csqrt_(cdiv(d.eps, d.mu));
    cdd minv = crcp_dd(m);
    const double zabs = cabs_(z);
    int nmx = nmx_rule(nmax, zabs) + (dn_extra < 0 ? 16 + 10 * (int)ceil(cbrt(zabs)) : dn_extra);
    cdd zinv = crcp_dd(z);
    const ddr xinv = rcp_dd(s);
    const int nsw = min((int)s, nmax);
    // D_n(mx) downward from D_nmx = 0 (mie_coeffs.py:118-121) and the psi ratio
    // psi_{n-1}/psi_n = D_n(x) + n/x above the oscillatory region: two
    // independent serial chains advanced in one loop so their latencies overlap.
    {
      const int sr = nsw < nmax ? ratio_start(nmax, s) : 0;   // ratio rows (nsw, sr]
      const int top = max(nmx - 1, sr);
      cd dd = cd{0.0, 0.0};
      double dr = 0.0;
      for (int k = top; k >= 1; --k) {
        const size_t o = (size_t)(k - 1) * 64 + lane;
        if (k <= nmx - 1) {
          const cd r = mul_dd((double)(k + 1), zinv);
          dd = r - crcp_fast(dd + r);
          if (k <= nmax) Dp[o] = dd;
        }
        if (k <= sr && k > nsw) {   // keeps psi_k / psi_{k-1} = 1/t
          const double r = mul_dd((double)k, xinv);
          const double it = rcp_nr(dr + r);
          if (k <= nmax) Rp[o] = it;
          dr = r - it;
        }
      }
    }
    double sn, cs;
    sincos(s, &sn, &cs);
    double psi_pp = cs, psi_p = sn;   // psi_{-1}, psi_0
    double chi_pp = -sn, chi_p = cs;  // chi_{-1}, chi_0
    Props pr;
    pr.init();
    const bool store = coef_off != nullptr;
    int64_t off = store ? coef_off[p] : 0;
    // one-row-ahead prefetch of the scratch rows (hides their latency
    // behind the previous row's arithmetic)
    cd D_nx = Dp[lane];
    double R_nx = Rp[lane];
#pragma unroll 2
    for (int n = 1; n <= nmax; ++n) {
      const cd D = D_nx;
      const double R = R_nx;
      if (n < nmax) {
        D_nx = Dp[(size_t)n * 64 + lane];
        R_nx = Rp[(size_t)n * 64 + lane];
      }
      double f = mul_dd((double)(2 * n - 1), xinv);
      double psi = (n <= nsw) ? fma(f, psi_p, -psi_pp) : psi_p * R;
      double chi = fma(f, chi_p, -chi_pp);
      double nx = mul_dd((double)n, xinv);
      cd da = mul_dd(D, minv) + nx;
      cd db = D * m + nx;
      cd xi = cd{psi, -chi};
      cd xi1 = cd{psi_p, -chi_p};
      cd a = cdiv(da * psi - psi_p, da * xi - xi1);
      cd b = cdiv(db * psi - psi_p, db * xi - xi1);
      pr.add(n, a, b);
      if (store) {
        an[2 * (off + n - 1)] = a.re;
        an[2 * (off + n - 1) + 1] = a.im;
        bn[2 * (off + n - 1)] = b.re;
        bn[2 * (off + n - 1) + 1] = b.im;
      }
      psi_pp = psi_p;
      psi_p = psi;
      chi_pp = chi_p;
      chi_p = chi;
    }
    if (props) pr.finish(s, p, n_all, props);
  }
}

// ---------------------------------------------------------------------------
// Homogeneous sphere, checkpointed downward chains
// ---------------------------------------------------------------------------
// The arithmetic of k_mie_homog with the memory schedule of
// k_mie_coated_ckpt: pass 1 runs D_n(mx) and the psi ratio chain but stores
// only their states every K rows (checkpoint b: the state after row bK+K+1;
// the ratio state only where block b has rows above the switch row), and
// rows K..1 straight to LDS; pass 2 walks the blocks upward, regenerating
// block b > 0 into LDS from checkpoint b (the next one in flight) with
// exactly pass 1's operations, so every row, and every result, is
// bit-identical to k_mie_homog's.  A lane-per-point step has little
// arithmetic to hide a scratch row's load behind (~115 FP64 instructions per
// row against ~24 B read and written per row and lane), so the full-row form
// is bound by its vector-memory instructions; here per-row traffic is one
// checkpoint per K rows and the upward steps read LDS.
// Slot layout: checkpoints [b][3][lane] (D.re, D.im, ratio state).
template <int MINW, int K>
__global__ __launch_bounds__(64, MINW) void k_mie_homog_ckpt(
    const int64_t* __restrict__ order, const int64_t* __restrict__ range, int64_t n_all,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ eps,
    const double* __restrict__ mu, const double* __restrict__ eps2,
    const int64_t* __restrict__ coef_off, double* __restrict__ an, double* __restrict__ bn,
    double* __restrict__ props, int32_t* __restrict__ status, double* __restrict__ scratch,
    int rows_cap, int dn_extra) {
  // rows above n_max in the last block stay below nmx - 1 (nmx >= n_max + 16)
  // and below the ratio chain's start (>= n_max + 18)
  static_assert(K >= 1 && K <= 12, "checkpoint interval");
  __shared__ double rows[K][3][64];
  const int lane = threadIdx.x;
  const int64_t begin = range[0], end = range[1];
  const int bcap = (rows_cap + K - 1) / K;
  double* CK = scratch + (size_t)blockIdx.x * (size_t)64 * 3 * bcap;   // [b][3][lane]

  for (int64_t base = begin + (int64_t)blockIdx.x * 64; base < end;
       base += (int64_t)gridDim.x * 64) {
    const int64_t idx = base + lane;
    if (idx >= end) continue;
    const int64_t p = order[idx];
    PointDesc d = describe(p, x, y, eps, mu, eps2);
    if (d.cls != 0) continue;
    const double s = d.size;
    const int nmax = nmax_rule(s);
    if (nmax > rows_cap) {
      atomicOr(status, 1);
      continue;
    }
    const cd z = csqrt_(d.eps * d.mu) * s;
    const cd m = csqrt_(cdiv(d.eps, d.mu));
    const cdd minv = crcp_dd(m);
    const double zabs = cabs_(z);
    const int nmx =
        nmx_rule(nmax, zabs) + (dn_extra < 0 ? 16 + 10 * (int)ceil(cbrt(zabs)) : dn_extra);
    const cdd zinv = crcp_dd(z);
    const ddr xinv = rcp_dd(s);
    // ratio rows (nsw, sr]; kNoRatio where x >= n_max (no ratio rows: the
    // upward form serves every n <= n_max)
    const int nsw = (int)s < nmax ? (int)s : kNoRatio;
    const int sr = nsw < nmax ? ratio_start(nmax, s) : 0;
    const int nblk = (nmax + K - 1) / K;
    // ---- pass 1: both downward chains, checkpointed every K rows
    {
      const int top = max(nmx - 1, sr);
      cd dd = cd{0.0, 0.0};
      double dr = 0.0;
      // rows above the ratio chain's start and the highest checkpoint (row
      // nblk K + 1) advance D_n alone; they exist only when top = nmx - 1
      const int lo = max(sr, nblk * K + 1);
      int k = top;
      for (; k > lo; --k) {
        const cd r = mul_dd((double)(k + 1), zinv);
        dd = r - crcp_fast(dd + r);
      }
      for (; k > K; --k) {
        if (k <= nmx - 1) {   // D_k = (k+1)/z - 1/(D_{k+1} + (k+1)/z), D_nmx = 0
          const cd r = mul_dd((double)(k + 1), zinv);
          dd = r - crcp_fast(dd + r);
        }
        if (k <= sr && k > nsw) {   // keeps psi_k / psi_{k-1} = 1/t
          const double r = mul_dd((double)k, xinv);
          dr = r - rcp_nr(dr + r);
        }
        const int b = (k - 1) / K - 1;      // state after row bK+K+1
        if ((k - 1) % K == 0 && b >= 1 && b < nblk) {
          double* c = CK + (size_t)b * 3 * 64 + lane;
          c[0] = dd.re;
          c[64] = dd.im;
          if (k - 1 > nsw) c[128] = dr;
        }
      }
      // rows K..1 (pass 2's first block) to LDS; k <= K < nmx - 1, and k <= sr
      // when the ratio chain is active
      for (int kk = min(top, K); kk >= 1; --kk) {
        {
          const cd r = mul_dd((double)(kk + 1), zinv);
          dd = r - crcp_fast(dd + r);
        }
        rows[kk - 1][0][lane] = dd.re;
        rows[kk - 1][1][lane] = dd.im;
        if (kk > nsw) {
          const double r = mul_dd((double)kk, xinv);
          const double it = rcp_nr(dr + r);
          rows[kk - 1][2][lane] = it;
          dr = r - it;
        }
      }
    }
    // ---- pass 2: per block, its K rows in LDS, then K upward steps
    double sn, cs;
    sincos(s, &sn, &cs);
    double psi_pp = cs, psi_p = sn;   // psi_{-1}, psi_0
    double chi_pp = -sn, chi_p = cs;  // chi_{-1}, chi_0
    Props pr;
    pr.init();
    const bool store = coef_off != nullptr;
    const int64_t off = store ? coef_off[p] : 0;
    double ck[3], ckn[3] = {0.0, 0.0, 0.0};
    auto load_ck = [&](int b) {
      const double* src = CK + (size_t)b * 3 * 64 + lane;
      ckn[0] = src[0];
      ckn[1] = src[64];
      if (b * K + K > nsw) ckn[2] = src[128];
    };
    if (nblk > 1) load_ck(1);
    for (int b = 0; b < nblk; ++b) {
      if (b > 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) ck[q] = ckn[q];
        if (b + 1 < nblk) load_ck(b + 1);
        cd dd = cd{ck[0], ck[1]};
        double dr = ck[2];
#pragma unroll
        for (int kk = K - 1; kk >= 0; --kk) {
          const int k = b * K + kk + 1;
          {
            const cd r = mul_dd((double)(k + 1), zinv);
            dd = r - crcp_fast(dd + r);
          }
          rows[kk][0][lane] = dd.re;
          rows[kk][1][lane] = dd.im;
          if (k > nsw) {
            const double r = mul_dd((double)k, xinv);
            const double it = rcp_nr(dr + r);
            rows[kk][2][lane] = it;
            dr = r - it;
          }
        }
      }
      const int kend = min(K, nmax - b * K);
#pragma unroll 2
      for (int kk = 0; kk < kend; ++kk) {
        const int n = b * K + kk + 1;
        const cd D = cd{rows[kk][0][lane], rows[kk][1][lane]};
        const double R = rows[kk][2][lane];
        double f = mul_dd((double)(2 * n - 1), xinv);
        double psi = (n <= nsw) ? fma(f, psi_p, -psi_pp) : psi_p * R;
        double chi = fma(f, chi_p, -chi_pp);
        double nx = mul_dd((double)n, xinv);
        cd da = mul_dd(D, minv) + nx;
        cd db = D * m + nx;
        cd xi = cd{psi, -chi};
        cd xi1 = cd{psi_p, -chi_p};
        cd a = cdiv(da * psi - psi_p, da * xi - xi1);
        cd bb = cdiv(db * psi - psi_p, db * xi - xi1);
        pr.add(n, a, bb);
        if (store) {
          an[2 * (off + n - 1)] = a.re;
          an[2 * (off + n - 1) + 1] = a.im;
          bn[2 * (off + n - 1)] = bb.re;
          bn[2 * (off + n - 1) + 1] = bb.im;
        }
        psi_pp = psi_p;
        psi_p = psi;
        chi_pp = chi_p;
        chi_p = chi;
      }
    }
    if (props) pr.finish(s, p, n_all, props);
  }
}

// ---------------------------------------------------------------------------
// Homogeneous sphere, P lanes per point (small batches)
// ---------------------------------------------------------------------------
// The lane-per-point kernel above needs thousands of points to fill the GPU,
// and its time is then the serial chain of the longest point.  Here each
// point's rows 1..nmax are split into P contiguous segments, one per lane of
// a P-lane group (64/P points per wavefront), and every serial chain of the
// algorithm is joined across the segments by a parallel scan of its
// transfer matrices:
//   * D_n downward (mie_coeffs.py:118-121): D_{k} = r - 1/(D_{k+1} + r),
//     r = (k+1)/z, is the Moebius map [[r, r^2-1], [1, r]] on (num, den);
//     the rows above nmax (down from the reference's start nmx, D = 0) are a
//     head split over the lanes as well.  A suffix scan of the segment
//     matrices gives each lane the D entering its segment from above, from
//     which it runs the recurrence over its own rows (the same operations as
//     the serial kernel) into its scratch rows.
//   * the psi ratio chain above n > x, the same form with r = k/x (real).
//   * psi_n (n <= x) and chi_n upward, psi_{n} = f psi_{n-1} - psi_{n-2},
//     f = (2n-1)/x: a prefix scan of [[f, -1], [1, 0]] gives each lane the
//     pair entering its segment.  Above x, psi_n = psi_x * prod R_k: the
//     psi at x comes from the lane that holds row x, the products of R from
//     a prefix-product scan.
// Each lane then forms a_n, b_n for its rows and partial efficiency sums,
// reduced over the group in a fixed order.  Matrix products are rescaled by
// powers of two (exact) to stay in range.  Against the serial kernel the
// results differ only by rounding (segment boundaries are reached by matrix
// products instead of the chain), checked against mpmath to x = 2094.
struct cm2 {
  cd a, b, c, d;   // [[a, b], [c, d]]
};
struct rm2 {
  double a, b, c, d;
};
__device__ __forceinline__ cm2 operator*(const cm2& u, const cm2& v) {
  return cm2{u.a * v.a + u.b * v.c, u.a * v.b + u.b * v.d, u.c * v.a + u.d * v.c,
             u.c * v.b + u.d * v.d};
}
__device__ __forceinline__ rm2 operator*(const rm2& u, const rm2& v) {
  return rm2{fma(u.a, v.a, u.b * v.c), fma(u.a, v.b, u.b * v.d), fma(u.c, v.a, u.d * v.c),
             fma(u.c, v.b, u.d * v.d)};
}
__device__ __forceinline__ cm2 cm2_eye() {
  return cm2{cd{1.0, 0.0}, cd{0.0, 0.0}, cd{0.0, 0.0}, cd{1.0, 0.0}};
}
__device__ __forceinline__ rm2 rm2_eye() { return rm2{1.0, 0.0, 0.0, 1.0}; }
// scale by a power of two so the largest component is in [0.5, 1) (exact)
__device__ __forceinline__ void renorm(cm2& m) {
  double mx = fmax(fmax(fmax(fabs(m.a.re), fabs(m.a.im)), fmax(fabs(m.b.re), fabs(m.b.im))),
                   fmax(fmax(fabs(m.c.re), fabs(m.c.im)), fmax(fabs(m.d.re), fabs(m.d.im))));
  if (mx == 0.0) return;
  const int e = __builtin_amdgcn_frexp_exp(mx);
  m.a = cd{ldexp(m.a.re, -e), ldexp(m.a.im, -e)};
  m.b = cd{ldexp(m.b.re, -e), ldexp(m.b.im, -e)};
  m.c = cd{ldexp(m.c.re, -e), ldexp(m.c.im, -e)};
  m.d = cd{ldexp(m.d.re, -e), ldexp(m.d.im, -e)};
}
__device__ __forceinline__ void renorm(rm2& m) {
  double mx = fmax(fmax(fabs(m.a), fabs(m.b)), fmax(fabs(m.c), fabs(m.d)));
  if (mx == 0.0) return;
  const int e = __builtin_amdgcn_frexp_exp(mx);
  m.a = ldexp(m.a, -e);
  m.b = ldexp(m.b, -e);
  m.c = ldexp(m.c, -e);
  m.d = ldexp(m.d, -e);
}
// shuffles within a group of P lanes (source lane index relative to the group)
template <int P>
__device__ __forceinline__ double grp_down(double v, int d) { return __shfl_down(v, d, P); }
template <int P>
__device__ __forceinline__ double grp_up(double v, int d) { return __shfl_up(v, d, P); }
template <int P>
__device__ __forceinline__ double grp_get(double v, int src) { return __shfl(v, src, P); }
template <int P>
__device__ __forceinline__ cd grp_down(cd v, int d) {
  return cd{grp_down<P>(v.re, d), grp_down<P>(v.im, d)};
}
template <int P>
__device__ __forceinline__ cm2 grp_down(const cm2& m, int d) {
  return cm2{grp_down<P>(m.a, d), grp_down<P>(m.b, d), grp_down<P>(m.c, d), grp_down<P>(m.d, d)};
}
template <int P>
__device__ __forceinline__ cm2 grp_get(const cm2& m, int src) {
  return cm2{cd{grp_get<P>(m.a.re, src), grp_get<P>(m.a.im, src)},
             cd{grp_get<P>(m.b.re, src), grp_get<P>(m.b.im, src)},
             cd{grp_get<P>(m.c.re, src), grp_get<P>(m.c.im, src)},
             cd{grp_get<P>(m.d.re, src), grp_get<P>(m.d.im, src)}};
}
template <int P>
__device__ __forceinline__ rm2 grp_down(const rm2& m, int d) {
  return rm2{grp_down<P>(m.a, d), grp_down<P>(m.b, d), grp_down<P>(m.c, d), grp_down<P>(m.d, d)};
}
template <int P>
__device__ __forceinline__ rm2 grp_up(const rm2& m, int d) {
  return rm2{grp_up<P>(m.a, d), grp_up<P>(m.b, d), grp_up<P>(m.c, d), grp_up<P>(m.d, d)};
}
template <int P>
__device__ __forceinline__ rm2 grp_get(const rm2& m, int src) {
  return rm2{grp_get<P>(m.a, src), grp_get<P>(m.b, src), grp_get<P>(m.c, src),
             grp_get<P>(m.d, src)};
}
// Inclusive suffix product over the group: lane j gets T_j T_{j+1} ... T_{P-1}.
template <int P, class M>
__device__ __forceinline__ M suffix_scan(M t, int j) {
#pragma unroll
  for (int d = 1; d < P; d <<= 1) {
    M o = grp_down<P>(t, d);
    if (j + d < P) {
      t = t * o;
      renorm(t);
    }
  }
  return t;
}
// Inclusive prefix product over the group: lane j gets U_j U_{j-1} ... U_0.
template <int P>
__device__ __forceinline__ rm2 prefix_scan(rm2 t, int j) {
#pragma unroll
  for (int d = 1; d < P; d <<= 1) {
    rm2 o = grp_up<P>(t, d);
    if (j >= d) t = t * o;
  }
  return t;
}
// sum over the group, the same butterfly order on every lane
template <int P>
__device__ __forceinline__ double grp_sum(double v) {
#pragma unroll
  for (int d = P / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, P);
  return v;
}

template <int P>
__global__ __launch_bounds__(64) void k_mie_homog_coop(
    const int64_t* __restrict__ order, const int64_t* __restrict__ range, int64_t n_all,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ eps,
    const double* __restrict__ mu, const double* __restrict__ eps2,
    const int64_t* __restrict__ coef_off, double* __restrict__ an, double* __restrict__ bn,
    double* __restrict__ props, int32_t* __restrict__ status, double* __restrict__ scratch,
    int lcap, int dn_extra) {
  constexpr int W = 64 / P;
  const int lane = threadIdx.x;
  const int j = lane % P, grp = lane / P;
  const int64_t begin = range[0], end = range[1];
  double* slot = scratch + (size_t)blockIdx.x * (size_t)lcap * 64 * 3;
  cd* Ds = reinterpret_cast<cd*>(slot);            // [row in segment][lane]
  double* Rs = slot + (size_t)lcap * 64 * 2;       // [row in segment][lane]

  for (int64_t base = begin + (int64_t)blockIdx.x * W; base < end;
       base += (int64_t)gridDim.x * W) {
    const int64_t idx = base + grp;
    if (idx >= end) continue;                      // the whole group: no shuffle partners lost
    const int64_t p = order[idx];
    PointDesc dsc = describe(p, x, y, eps, mu, eps2);
    if (dsc.cls != 0) continue;
    const double s = dsc.size;
    const int nmax = nmax_rule(s);
    const int L = (nmax + P - 1) / P;
    if (L > lcap) {
      if (j == 0) atomicOr(status, 1);
      continue;
    }
    const cd z = csqrt_(dsc.eps * dsc.mu) * s;
    const cd m = csqrt_(cdiv(dsc.eps, dsc.mu));
    const cdd minv = crcp_dd(m);
    const double zabs = cabs_(z);
    const int nmx =
        nmx_rule(nmax, zabs) + (dn_extra < 0 ? 16 + 10 * (int)ceil(cbrt(zabs)) : dn_extra);
    const cdd zinv = crcp_dd(z);
    const ddr xinv = rcp_dd(s);
    const int nsw = min((int)s, nmax);
    const int n0 = j * L + 1, n1 = min((j + 1) * L, nmax);   // this lane's rows (may be empty)

    // ---- D_n: head rows nmax+1 .. nmx-1, then this lane's rows
    cd d_in;
    {
      const int H = nmx - 1 - nmax;
      const int HL = (H + P - 1) / P;
      const int h0 = nmax + 1 + j * HL, h1 = min(nmax + (j + 1) * HL, nmx - 1);
      cm2 t = cm2_eye();
      int cnt = 0;
      for (int k = h1; k >= h0; --k) {
        const cd r = mul_dd((double)(k + 1), zinv);
        const cm2 mk{r, r * r - 1.0, cd{1.0, 0.0}, r};
        t = mk * t;
        if ((++cnt & 3) == 0) renorm(t);
      }
      renorm(t);
      const cm2 head = grp_get<P>(suffix_scan<P>(t, j), 0);
      const cd vh_a = head.b, vh_b = head.d;     // head (0, 1)
      t = cm2_eye();
      cnt = 0;
      for (int k = n1; k >= n0; --k) {
        const cd r = mul_dd((double)(k + 1), zinv);
        const cm2 mk{r, r * r - 1.0, cd{1.0, 0.0}, r};
        t = mk * t;
        if ((++cnt & 3) == 0) renorm(t);
      }
      renorm(t);
      cm2 sfx = grp_down<P>(suffix_scan<P>(t, j), 1);   // T_{j+1} ... T_{P-1}
      if (j == P - 1) sfx = cm2_eye();
      const cd va = sfx.a * vh_a + sfx.b * vh_b, vb = sfx.c * vh_a + sfx.d * vh_b;
      d_in = cdiv(va, vb);
    }
    {
      cd dd = d_in;
      for (int k = n1; k >= n0; --k) {
        const cd r = mul_dd((double)(k + 1), zinv);
        dd = r - crcp_fast(dd + r);
        Ds[(size_t)(k - n0) * 64 + lane] = dd;
      }
    }
    // ---- psi ratios above nsw: head nmax+1 .. sr, then this lane's rows > nsw
    double seg_prod = 1.0;
    if (nsw < nmax) {
      const int sr = ratio_start(nmax, s);
      const int H = sr - nmax;
      const int HL = (H + P - 1) / P;
      const int h0 = nmax + 1 + j * HL, h1 = min(nmax + (j + 1) * HL, sr);
      rm2 t = rm2_eye();
      for (int k = h1; k >= h0; --k) {
        const double r = mul_dd((double)k, xinv);
        t = rm2{r, fma(r, r, -1.0), 1.0, r} * t;
        renorm(t);
      }
      const rm2 head = grp_get<P>(suffix_scan<P>(t, j), 0);
      const double vh_a = head.b, vh_b = head.d;
      t = rm2_eye();
      const int lo = max(n0, nsw + 1);
      for (int k = n1; k >= lo; --k) {
        const double r = mul_dd((double)k, xinv);
        t = rm2{r, fma(r, r, -1.0), 1.0, r} * t;
        renorm(t);
      }
      rm2 sfx = grp_down<P>(suffix_scan<P>(t, j), 1);
      if (j == P - 1) sfx = rm2_eye();
      double dr = fma(sfx.a, vh_a, sfx.b * vh_b) / fma(sfx.c, vh_a, sfx.d * vh_b);
      for (int k = n1; k >= lo; --k) {
        const double r = mul_dd((double)k, xinv);
        const double it = rcp_nr(dr + r);
        Rs[(size_t)(k - n0) * 64 + lane] = it;
        seg_prod *= it;
        dr = r - it;
      }
    }
    // ---- psi (n <= nsw) and chi upward: the pair entering this lane's rows
    double sn, cs;
    sincos(s, &sn, &cs);
    double psi_p, psi_pp, chi_p, chi_pp;
    {
      rm2 u = rm2_eye();
      for (int n = n0; n <= n1; ++n) {
        const double f = mul_dd((double)(2 * n - 1), xinv);
        u = rm2{f, -1.0, 1.0, 0.0} * u;
      }
      rm2 e = prefix_scan<P>(u, j);
      rm2 ex = grp_up<P>(e, 1);                    // U_{j-1} ... U_0
      if (j == 0) ex = rm2_eye();
      psi_p = fma(ex.a, sn, ex.b * cs);            // psi_{n0-1}
      psi_pp = fma(ex.c, sn, ex.d * cs);
      chi_p = fma(ex.a, cs, -(ex.b * sn));
      chi_pp = fma(ex.c, cs, -(ex.d * sn));
    }
    if (nsw < nmax) {
      // psi at nsw, from the lane whose rows hold it, upward from its pair
      double pn = sn;
      const int js = nsw >= 1 ? (nsw - 1) / L : 0;
      if (nsw >= 1 && j == js) {
        double a = psi_p, b = psi_pp;
        for (int n = n0; n <= nsw; ++n) {
          const double f = mul_dd((double)(2 * n - 1), xinv);
          const double c = fma(f, a, -b);
          b = a;
          a = c;
        }
        pn = a;
      }
      if (nsw >= 1) pn = __shfl(pn, js, P);
      // prefix product of the ratios below this lane's rows
      double pre = seg_prod;
#pragma unroll
      for (int d = 1; d < P; d <<= 1) {
        const double o = __shfl_up(pre, d, P);
        if (j >= d) pre *= o;
      }
      pre = __shfl_up(pre, 1, P);
      if (j == 0) pre = 1.0;
      if (n0 - 1 >= nsw) psi_p = pn * pre;
    }
    // ---- a_n, b_n and partial efficiency sums over this lane's rows
    const bool store = coef_off != nullptr;
    const int64_t off = store ? coef_off[p] : 0;
    double sext = 0.0, ssca = 0.0, asy1 = 0.0, asy2 = 0.0;
    cd qb = cd{0.0, 0.0}, a_first = cd{0.0, 0.0}, b_first = cd{0.0, 0.0};
    cd ap = cd{0.0, 0.0}, bp = cd{0.0, 0.0};
    for (int n = n0; n <= n1; ++n) {
      const cd D = Ds[(size_t)(n - n0) * 64 + lane];
      const double f = mul_dd((double)(2 * n - 1), xinv);
      const double psi = (n <= nsw) ? fma(f, psi_p, -psi_pp)
                                    : psi_p * Rs[(size_t)(n - n0) * 64 + lane];
      const double chi = fma(f, chi_p, -chi_pp);
      const double nx = mul_dd((double)n, xinv);
      const cd da = mul_dd(D, minv) + nx;
      const cd db = D * m + nx;
      const cd xi = cd{psi, -chi};
      const cd xi1 = cd{psi_p, -chi_p};
      const cd a = cdiv(da * psi - psi_p, da * xi - xi1);
      const cd b = cdiv(db * psi - psi_p, db * xi - xi1);
      const double dn = (double)n, cn = 2.0 * dn + 1.0;
      sext = fma(cn, a.re + b.re, sext);
      ssca = fma(cn, norm2(a) + norm2(b), ssca);
      qb = qb + (a - b) * ((n & 1) ? -cn : cn);
      if (n > n0) {
        const double dm = dn - 1.0;
        asy1 = fma(dm * (dm + 2.0) * rcp_nr(dm + 1.0),
                   ap.re * a.re + ap.im * a.im + bp.re * b.re + bp.im * b.im, asy1);
      } else {
        a_first = a;
        b_first = b;
      }
      asy2 = fma(cn * rcp_nr(dn * (dn + 1.0)), a.re * b.re + a.im * b.im, asy2);
      ap = a;
      bp = b;
      if (store) {
        an[2 * (off + n - 1)] = a.re;
        an[2 * (off + n - 1) + 1] = a.im;
        bn[2 * (off + n - 1)] = b.re;
        bn[2 * (off + n - 1) + 1] = b.im;
      }
      psi_pp = psi_p;
      psi_p = psi;
      chi_pp = chi_p;
      chi_p = chi;
    }
    // the pair (n1, n1 + 1) across the segment boundary
    {
      const cd an1 = grp_down<P>(a_first, 1), bn1 = grp_down<P>(b_first, 1);
      if (j < P - 1 && n1 >= n0 && n1 + 1 <= nmax) {
        const double dm = (double)n1;
        asy1 = fma(dm * (dm + 2.0) * rcp_nr(dm + 1.0),
                   ap.re * an1.re + ap.im * an1.im + bp.re * bn1.re + bp.im * bn1.im, asy1);
      }
    }
    sext = grp_sum<P>(sext);
    ssca = grp_sum<P>(ssca);
    asy1 = grp_sum<P>(asy1);
    asy2 = grp_sum<P>(asy2);
    qb = cd{grp_sum<P>(qb.re), grp_sum<P>(qb.im)};
    if (j == 0 && props) {
      const double y2 = s * s;
      const double qext = 2.0 * sext / y2, qsca = 2.0 * ssca / y2;
      const double qbv = (qb.re * qb.re + qb.im * qb.im) / y2;
      props[0 * n_all + p] = qext;
      props[1 * n_all + p] = qsca;
      props[2 * n_all + p] = qext - qsca;
      props[3 * n_all + p] = qbv;
      props[4 * n_all + p] = 4.0 / y2 * (asy1 + asy2) / qsca;
      props[5 * n_all + p] = qbv / qsca;
    }
  }
}

// ---------------------------------------------------------------------------
// Recurrence helpers (one lane, rows [n-1][lane] of a scratch plane)
// ---------------------------------------------------------------------------
// D_1..D_nmax of z from D_nstart = 0 (mie_coeffs.py:118-121); subtract=true
// stores plane[row] - D instead (difference of two starts).
__device__ __forceinline__ void dn_chain(cdd zi, int nstart, int nmax, cd* __restrict__ plane,
                                         int lane, bool subtract) {
  cd dd = cd{0.0, 0.0};
  for (int j = nstart - 1; j >= 1; --j) {
    cd r = mul_dd((double)(j + 1), zi);
    dd = r - crcp_fast(dd + r);
    if (j <= nmax) {
      const size_t o = (size_t)(j - 1) * 64 + lane;
      plane[o] = subtract ? plane[o] - dd : dd;
    }
  }
}
// ---------------------------------------------------------------------------
// Range-checked loads from a block's scratch slot
// ---------------------------------------------------------------------------
// A raw buffer load whose offset lies past the resource's byte range returns
// zeros without touching memory.  A row that only some lanes need (a ratio
// row, a start-index correction, the next row past nmax) is thus loaded
// without a branch: the other lanes pass kNoRow.  Slots are < 2 GiB (checked
// on the host), so kNoRow is always out of range.
constexpr uint32_t kNoRow = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slot_rsrc(double* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ cd ld_cd(__amdgpu_buffer_rsrc_t r, uint32_t off, bool ok) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, ok ? off : kNoRow, 0, 0);
  return cd{__builtin_bit_cast(double, (unsigned long long)v[0] | ((unsigned long long)v[1] << 32)),
            __builtin_bit_cast(double, (unsigned long long)v[2] | ((unsigned long long)v[3] << 32))};
}
__device__ __forceinline__ double ld_d(__amdgpu_buffer_rsrc_t r, uint32_t off, bool ok) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, ok ? off : kNoRow, 0, 0);
  return __builtin_bit_cast(double, (unsigned long long)v[0] | ((unsigned long long)v[1] << 32));
}

// ---------------------------------------------------------------------------
// Coated sphere, real or weakly absorbing shell (|Im m2| y <= 1)
// ---------------------------------------------------------------------------
// The coefficients of mie_coeffs.py:183-251, with the shell combination
// (:230-244) rearranged so that nothing is divided by psi_n(m2 y) and no
// D_n(m2 x) is multiplied by psi_n(m2 x).  The literal formula does both: near
// a zero of either function its FP64 evaluation (the reference's included)
// loses up to -2 log10|psi| digits, up to 1e-3 in Q.  With psi'_n =
// psi_{n-1} - (n/z) psi_n, chi' likewise and psi' chi - psi chi' = 1:
//   uu psi_v = m D_u psi_v - psi'_v - dD_v psi_v
//   dns      = [psi'_w + uu psi_v (psi'_w chi_v - chi'_w psi_v)]
//            / [psi_w  + uu psi_v (psi_w  chi_v - chi_w  psi_v)] + dD_w
// (vv and gns alike), identical in exact arithmetic to the reference's
// dns1 + dnw.  dD = D_ref - D_exact corrects for the reference's start index
// (D_ref from nmx = round(max(nmax, |m1 y|, |w|) + 16)); it vanishes unless
// that start is within a few Airy widths of |v| or |w|, and is only then
// computed (two extra recurrences).  D_n(m2 x) and D_n(m2 y) are otherwise
// not needed at all: pass 1 runs only the D_n(m1 x) recurrence.
// One point of pass 2 (upward recurrences of psi, chi at v = m2 x, w = m2 y
// and y, the shell combination, a_n, b_n and the efficiency sums), on rows
// stored by pass 1 in the slot [DU | RV | RW | DDV | DDW | RY] (planes of
// `plane` doubles; complex planes take two).  Every row's arithmetic is
// branch-free: the upward and the ratio form of psi are both formed and one
// is selected, and rows only some lanes need are loaded through the range
// check (zeros elsewhere, which leave the results unchanged: U - 0 * psi = U,
// dns + 0 = dns).  DD: the start-index corrections; STORE: write a_n, b_n.
struct CoatedPt {
  cd v, w;
  double y;
  cd m, m2;
  int nmax, nsw_v, nsw_w, nsw_y;
  bool ddv_on, ddw_on;
};
template <bool DD, bool STORE>
__device__ __forceinline__ void coated_up(const CoatedPt& c, __amdgpu_buffer_rsrc_t rs,
                                          uint32_t plane, int lane, int64_t off,
                                          double* __restrict__ an, double* __restrict__ bn,
                                          Props& pr) {
  const cdd vinv = crcp_dd(c.v), winv = crcp_dd(c.w), minv = crcp_dd(c.m);
  const ddr yinv = rcp_dd(c.y);
  const cd m = c.m, m2 = c.m2;
  const int nmax = c.nmax, nsw_v = c.nsw_v, nsw_w = c.nsw_w, nsw_y = c.nsw_y;
  cd pv_p, pv_pp;   // psi_0(v) = sin v, psi_{-1}(v) = cos v
  csincos_(c.v, pv_p, pv_pp);
  cd cv_pp = -pv_p, cv_p = pv_pp;             // chi_{-1}(v), chi_0(v)
  cd pw_p, pw_pp;
  csincos_(c.w, pw_p, pw_pp);
  cd cw_pp = -pw_p, cw_p = pw_pp;
  double sn, cs;
  sincos(c.y, &sn, &cs);
  double py_pp = cs, py_p = sn, cy_pp = -sn, cy_p = cs;
  // byte offsets of row 1 of each plane for this lane; a row adds 64 entries
  const uint32_t b_du = 16u * (uint32_t)lane, b_rv = 16u * (plane + (uint32_t)lane),
                 b_rw = 16u * (2u * plane + (uint32_t)lane),
                 b_ddv = 16u * (3u * plane + (uint32_t)lane),
                 b_ddw = 16u * (4u * plane + (uint32_t)lane),
                 b_ry = 8u * (10u * plane + (uint32_t)lane);
  // the stored rows of step n + 1 are loaded during step n
  cd du_nx = ld_cd(rs, b_du, true), rv_nx = ld_cd(rs, b_rv, nsw_v < 1),
     rw_nx = ld_cd(rs, b_rw, nsw_w < 1);
  double ry_nx = ld_d(rs, b_ry, nsw_y < 1);
#pragma unroll 2
  for (int n = 1; n <= nmax; ++n) {
    const uint32_t r = (uint32_t)(n - 1) * 64u;
    const cd du = du_nx, rv = rv_nx, rw = rw_nx;
    const double ry = ry_nx;
    const bool more = n < nmax;
    du_nx = ld_cd(rs, b_du + 16u * (r + 64u), more);
    rv_nx = ld_cd(rs, b_rv + 16u * (r + 64u), more && n >= nsw_v);
    rw_nx = ld_cd(rs, b_rw + 16u * (r + 64u), more && n >= nsw_w);
    ry_nx = ld_d(rs, b_ry + 8u * (r + 64u), more && n >= nsw_y);
    const double fn = (double)(2 * n - 1), dn = (double)n;
    const cd fv = mul_dd(fn, vinv), fw = mul_dd(fn, winv);
    const cd pv_up = fv * pv_p - pv_pp, pv_rt = pv_p * rv;
    const cd pv = (n <= nsw_v) ? pv_up : pv_rt;
    const cd chv = fv * cv_p - cv_pp;
    const cd pw_up = fw * pw_p - pw_pp, pw_rt = pw_p * rw;
    const cd pw = (n <= nsw_w) ? pw_up : pw_rt;
    const cd chw = fw * cw_p - cw_pp;
    const double f = mul_dd(fn, yinv);
    const double py_up = fma(f, py_p, -py_pp), py_rt = py_p * ry;
    const double py = (n <= nsw_y) ? py_up : py_rt;
    const double chy = fma(f, cy_p, -cy_pp);
    // derivatives psi'_n = psi_{n-1} - (n/z) psi_n
    const cd dpv = pv_p - mul_dd(dn, vinv) * pv;
    const cd nw = mul_dd(dn, winv);
    const cd dpw = pw_p - nw * pw, dcw = cw_p - nw * chw;
    const cd t = du * pv;
    cd U = m * t - dpv, V = mul_dd(t, minv) - dpv;
    if constexpr (DD) {
      const cd cc = ld_cd(rs, b_ddv + 16u * r, c.ddv_on) * pv;
      U = U - cc;
      V = V - cc;
    }
    const cd A = pw * chv - chw * pv, B = dpw * chv - dcw * pv;
    // dns = Nd / Dd and gns = Ng / Dg (plus dD_w: N += dD_w D)
    cd Nd = dpw + U * B, Ng = dpw + V * B;
    const cd Dd = pw + U * A, Dg = pw + V * A;
    if constexpr (DD) {
      const cd cc = ld_cd(rs, b_ddw + 16u * r, c.ddw_on);
      Nd = Nd + cc * Dd;
      Ng = Ng + cc * Dg;
    }
    const double nrat = mul_dd(dn, yinv);
    const cd xi = cd{py, -chy}, xi1 = cd{py_p, -cy_p};
    const cd Ma = m2 * Dd, Pa = Nd + Ma * nrat, Pb = m2 * Ng + Dg * nrat;
    cd a, b;
    cdiv2(Pa * py - Ma * py_p, Pa * xi - Ma * xi1, Pb * py - Dg * py_p, Pb * xi - Dg * xi1, a, b);
    pr.add_seq(n, a, b);
    if constexpr (STORE) {
      an[2 * (off + n - 1)] = a.re;
      an[2 * (off + n - 1) + 1] = a.im;
      bn[2 * (off + n - 1)] = b.re;
      bn[2 * (off + n - 1) + 1] = b.im;
    }
    pv_pp = pv_p; pv_p = pv; cv_pp = cv_p; cv_p = chv;
    pw_pp = pw_p; pw_p = pw; cw_pp = cw_p; cw_p = chw;
    py_pp = py_p; py_p = py; cy_pp = cy_p; cy_p = chy;
  }
}

template <int MINW>
__global__ __launch_bounds__(64, MINW) void k_mie_coated_stable(
    const int64_t* __restrict__ order, const int64_t* __restrict__ range, int64_t n_all,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ eps,
    const double* __restrict__ mu, const double* __restrict__ eps2,
    const int64_t* __restrict__ coef_off, double* __restrict__ an, double* __restrict__ bn,
    double* __restrict__ props, int32_t* __restrict__ status, double* __restrict__ scratch,
    int rows_cap, int dn_extra) {
  const int lane = threadIdx.x;
  const int64_t begin = range[2], end = range[3];
  // 11 doubles per row: DU, RV, RW, DDV, DDW (complex), RY (real)
  double* slot = scratch + (size_t)blockIdx.x * (size_t)rows_cap * 64 * 11;
  const size_t plane = (size_t)rows_cap * 64;
  cd* DU = reinterpret_cast<cd*>(slot);
  cd* RV = reinterpret_cast<cd*>(slot + 2 * plane);
  cd* RW = reinterpret_cast<cd*>(slot + 4 * plane);
  cd* DDV = reinterpret_cast<cd*>(slot + 6 * plane);
  cd* DDW = reinterpret_cast<cd*>(slot + 8 * plane);
  double* RY = slot + 10 * plane;

  for (int64_t base = begin + (int64_t)blockIdx.x * 64; base < end;
       base += (int64_t)gridDim.x * 64) {
    const int64_t idx = base + lane;
    if (idx >= end) continue;
    const int64_t p = order[idx];
    PointDesc d = describe(p, x, y, eps, mu, eps2);
    if (d.cls != 1) continue;
    const double yy = d.size, xx = d.xcore;
    const int nmax = nmax_rule(yy);
    if (nmax > rows_cap) {
      atomicOr(status, 1);
      continue;
    }
    const cd m1 = csqrt_(d.eps), m2 = csqrt_(d.eps2);
    const cd m = cdiv(m2, m1);
    const cd u = m1 * xx, v = m2 * xx, w = m2 * yy;
    const double mxv = fmax(cabs_(m1 * yy), cabs_(w));
    const int nmx =
        nmx_rule(nmax, mxv) + (dn_extra < 0 ? 16 + 10 * (int)ceil(cbrt(mxv)) : dn_extra);
    const double av = cabs_(v), aw = cabs_(w);
    // ---- pass 1: downward recurrences.  D_n(u) and the psi ratios of v, w, y
    // are four independent serial chains; one loop advances all of them so
    // their dependent FP64 latencies overlap (each chain's own sequence of
    // operations is unchanged).
    const int nsw_v = min((int)av, nmax), nsw_w = min((int)aw, nmax), nsw_y = min((int)yy, nmax);
    {
      const cdd ui = crcp_dd(u), vi = crcp_dd(v), wi = crcp_dd(w);
      const ddr yi = rcp_dd(yy);
      const int sv = nsw_v < nmax ? ratio_start(nmax, av) : 0;   // ratio rows (nsw, s]
      const int sw = nsw_w < nmax ? ratio_start(nmax, aw) : 0;
      const int sy = nsw_y < nmax ? ratio_start(nmax, yy) : 0;
      const int top = max(max(nmx - 1, sv), max(sw, sy));
      cd du = cd{0.0, 0.0}, dv = cd{0.0, 0.0}, dw = cd{0.0, 0.0};
      double dy = 0.0;
      for (int k = top; k >= 1; --k) {
        const size_t o = (size_t)(k - 1) * 64 + lane;
        if (k <= nmx - 1) {   // D_k(u) = (k+1)/u - 1/(D_{k+1} + (k+1)/u), D_nmx = 0
          const cd r = mul_dd((double)(k + 1), ui);
          du = r - crcp_fast(du + r);
          if (k <= nmax) DU[o] = du;
        }
        // ratio chains keep psi_k / psi_{k-1} = 1/t (t = psi_{k-1}/psi_k), which
        // the chain forms anyway; the upward pass then multiplies
        if (k <= sv && k > nsw_v) {
          const cd r = mul_dd((double)k, vi);
          const cd it = crcp_fast(dv + r);
          if (k <= nmax) RV[o] = it;
          dv = r - it;
        }
        if (k <= sw && k > nsw_w) {
          const cd r = mul_dd((double)k, wi);
          const cd it = crcp_fast(dw + r);
          if (k <= nmax) RW[o] = it;
          dw = r - it;
        }
        if (k <= sy && k > nsw_y) {
          const double r = mul_dd((double)k, yi);
          const double it = rcp_nr(dy + r);
          if (k <= nmax) RY[o] = it;
          dy = r - it;
        }
      }
    }
    const bool ddv_on = !converged_start(nmx, av), ddw_on = !converged_start(nmx, aw);
    if (ddv_on) {
      const cdd vi = crcp_dd(v);
      dn_chain(vi, nmx, nmax, DDV, lane, false);
      dn_chain(vi, nmx + 16 + 8 * (int)ceil(cbrt(av)), nmax, DDV, lane, true);
    }
    if (ddw_on) {
      const cdd wi = crcp_dd(w);
      dn_chain(wi, nmx, nmax, DDW, lane, false);
      dn_chain(wi, nmx + 16 + 8 * (int)ceil(cbrt(aw)), nmax, DDW, lane, true);
    }
    // ---- pass 2: upward Riccati-Bessel recurrences and the combination,
    // without divergent branches (coated_up); the start-index corrections
    // are compiled in only when a lane of the wavefront needs them
    Props pr;
    pr.init_seq();
    const int64_t off = coef_off ? coef_off[p] : 0;
    const CoatedPt cp{v, w, yy, m, m2, nmax, nsw_v, nsw_w, nsw_y, ddv_on, ddw_on};
    const __amdgpu_buffer_rsrc_t rs = slot_rsrc(slot, (uint32_t)(plane * 11 * sizeof(double)));
    if (__any(ddv_on || ddw_on)) {
      if (coef_off) coated_up<true, true>(cp, rs, (uint32_t)plane, lane, off, an, bn, pr);
      else coated_up<true, false>(cp, rs, (uint32_t)plane, lane, off, an, bn, pr);
    } else {
      if (coef_off) coated_up<false, true>(cp, rs, (uint32_t)plane, lane, off, an, bn, pr);
      else coated_up<false, false>(cp, rs, (uint32_t)plane, lane, off, an, bn, pr);
    }
    if (props) pr.finish(yy, p, n_all, props);
  }
}

// ---------------------------------------------------------------------------
// Coated sphere, real or weakly absorbing shell, checkpointed downward chains
// ---------------------------------------------------------------------------
// Pass 2 of k_mie_coated_ckpt for one point: per block, its K rows in LDS
// (block 0 as pass 1 left them; block b > 0 regenerated from checkpoint b,
// the next one in flight, with exactly pass 1's operations), then the K
// upward steps, unrolled by two.  The upward steps are branch-free as in
// coated_up (both psi forms formed, one selected; the LDS rows of an
// inactive chain are read and ignored), the start-index corrections come
// through range-checked loads and are compiled in only when a lane of the
// wavefront needs them (DD); STORE writes a_n, b_n.
// nsw_v, nsw_w, nsw_y: the last row of the upward psi form; a chain with no
// ratio rows (|z| >= n_max) carries kNoRatio there, so "row k uses the
// ratio chain" is k > nsw alone.  Pass 2 never reaches a row above
// nblk K <= n_max + K - 1, which lies below nmx - 1 (nmx >= n_max + 16) and
// below every active chain's start (>= n_max + 18): the regeneration needs
// neither bound.
struct CkptPt {
  cd v, w, m, m2;
  double y;
  cdd ui, vi, wi, minv;
  ddr yi;
  int nmax, nsw_v, nsw_w, nsw_y, nblk;
  bool ddv_on, ddw_on;
};
template <int K, bool DD, bool STORE>
__device__ __forceinline__ Props ckpt_pass2(const CkptPt& c, double (*rows)[7][64],
                                            const double* __restrict__ CK,
                                            __amdgpu_buffer_rsrc_t rs, uint32_t b_ddv,
                                            uint32_t b_ddw, int lane, int64_t off,
                                            double* __restrict__ an, double* __restrict__ bn) {
  cd pv_p, pv_pp;   // psi_0(v) = sin v, psi_{-1}(v) = cos v
  csincos_(c.v, pv_p, pv_pp);
  cd cv_pp = -pv_p, cv_p = pv_pp;             // chi_{-1}(v), chi_0(v)
  cd pw_p, pw_pp;
  csincos_(c.w, pw_p, pw_pp);
  cd cw_pp = -pw_p, cw_p = pw_pp;
  double sn, cs;
  sincos(c.y, &sn, &cs);
  double py_pp = cs, py_p = sn, cy_pp = -sn, cy_p = cs;
  const cd m = c.m, m2 = c.m2;
  Props pr;
  pr.init_seq();
  // block 0's rows are in LDS already (pass 1 leaves them there); block b > 0
  // is regenerated from checkpoint b, the next checkpoint in flight meanwhile
  // checkpoint b holds a ratio chain's state only where block b uses it
  // (its top row bK + K above the chain's nsw); the other words are left
  // unloaded (keeping the earlier block's), and the regeneration never uses
  // them
  double ck[7], ckn[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  auto load_ck = [&](int b) {
    const double* src = CK + (size_t)b * 7 * 64 + lane;
    const int top = b * K + K;
    ckn[0] = src[0];
    ckn[1] = src[64];
    if (top > c.nsw_v) {
      ckn[2] = src[128];
      ckn[3] = src[192];
    }
    if (top > c.nsw_w) {
      ckn[4] = src[256];
      ckn[5] = src[320];
    }
    if (top > c.nsw_y) ckn[6] = src[384];
  };
  if (c.nblk > 1) load_ck(1);
  for (int b = 0; b < c.nblk; ++b) {
    if (b > 0) {
#pragma unroll
      for (int q = 0; q < 7; ++q) ck[q] = ckn[q];
      if (b + 1 < c.nblk) load_ck(b + 1);
      cd du = cd{ck[0], ck[1]}, dv = cd{ck[2], ck[3]}, dw = cd{ck[4], ck[5]};
      double dy = ck[6];
#pragma unroll
      for (int kk = K - 1; kk >= 0; --kk) {
        const int k = b * K + kk + 1;
        {
          const cd r = mul_dd((double)(k + 1), c.ui);
          du = r - crcp_fast(du + r);
        }
        rows[kk][0][lane] = du.re;
        rows[kk][1][lane] = du.im;
        if (k > c.nsw_v) {
          const cd r = mul_dd((double)k, c.vi);
          const cd it = crcp_fast(dv + r);
          rows[kk][2][lane] = it.re;
          rows[kk][3][lane] = it.im;
          dv = r - it;
        }
        if (k > c.nsw_w) {
          const cd r = mul_dd((double)k, c.wi);
          const cd it = crcp_fast(dw + r);
          rows[kk][4][lane] = it.re;
          rows[kk][5][lane] = it.im;
          dw = r - it;
        }
        if (k > c.nsw_y) {
          const double r = mul_dd((double)k, c.yi);
          const double it = rcp_nr(dy + r);
          rows[kk][6][lane] = it;
          dy = r - it;
        }
      }
    }
    const int kend = min(K, c.nmax - b * K);
    // unrolled by two: the psi/chi state rotates through registers without moves
#pragma unroll 2
    for (int kk = 0; kk < kend; ++kk) {
      const int n = b * K + kk + 1;
      const uint32_t r = (uint32_t)(n - 1) * 64u;
      const cd du = cd{rows[kk][0][lane], rows[kk][1][lane]};
      const cd rv = cd{rows[kk][2][lane], rows[kk][3][lane]};
      const cd rw = cd{rows[kk][4][lane], rows[kk][5][lane]};
      const double ry = rows[kk][6][lane];
      const double fn = (double)(2 * n - 1), dn = (double)n;
      const cd fv = mul_dd(fn, c.vi), fw = mul_dd(fn, c.wi);
      const cd pv_up = fv * pv_p - pv_pp, pv_rt = pv_p * rv;
      const cd pv = (n <= c.nsw_v) ? pv_up : pv_rt;
      const cd chv = fv * cv_p - cv_pp;
      const cd pw_up = fw * pw_p - pw_pp, pw_rt = pw_p * rw;
      const cd pw = (n <= c.nsw_w) ? pw_up : pw_rt;
      const cd chw = fw * cw_p - cw_pp;
      const double f = mul_dd(fn, c.yi);
      const double py_up = fma(f, py_p, -py_pp), py_rt = py_p * ry;
      const double py = (n <= c.nsw_y) ? py_up : py_rt;
      const double chy = fma(f, cy_p, -cy_pp);
      const cd dpv = pv_p - mul_dd(dn, c.vi) * pv;
      const cd nw = mul_dd(dn, c.wi);
      const cd dpw = pw_p - nw * pw, dcw = cw_p - nw * chw;
      const cd t = du * pv;
      cd U = m * t - dpv, V = mul_dd(t, c.minv) - dpv;
      if constexpr (DD) {
        const cd cc = ld_cd(rs, b_ddv + 16u * r, c.ddv_on) * pv;
        U = U - cc;
        V = V - cc;
      }
      const cd A = pw * chv - chw * pv, B = dpw * chv - dcw * pv;
      cd Nd = dpw + U * B, Ng = dpw + V * B;
      const cd Dd = pw + U * A, Dg = pw + V * A;
      if constexpr (DD) {
        const cd cc = ld_cd(rs, b_ddw + 16u * r, c.ddw_on);
        Nd = Nd + cc * Dd;
        Ng = Ng + cc * Dg;
      }
      const double nrat = mul_dd(dn, c.yi);
      const cd xi = cd{py, -chy}, xi1 = cd{py_p, -cy_p};
      const cd Ma = m2 * Dd, Pa = Nd + Ma * nrat, Pb = m2 * Ng + Dg * nrat;
      cd a, bb;
      cdiv2(Pa * py - Ma * py_p, Pa * xi - Ma * xi1, Pb * py - Dg * py_p, Pb * xi - Dg * xi1, a,
            bb);
      pr.add_seq(n, a, bb);
      if constexpr (STORE) {
        an[2 * (off + n - 1)] = a.re;
        an[2 * (off + n - 1) + 1] = a.im;
        bn[2 * (off + n - 1)] = bb.re;
        bn[2 * (off + n - 1) + 1] = bb.im;
      }
      pv_pp = pv_p; pv_p = pv; cv_pp = cv_p; cv_p = chv;
      pw_pp = pw_p; pw_p = pw; cw_pp = cw_p; cw_p = chw;
      py_pp = py_p; py_p = py; cy_pp = cy_p; cy_p = chy;
    }
  }
  return pr;
}

// The arithmetic of k_mie_coated_stable, with a different memory schedule.
// Pass 1 runs the four downward chains (D_n(u) and the psi ratios of v, w,
// y) over every row but stores only their states every K rows: checkpoint b
// is the state after row bK+K+1, from which rows bK+K .. bK+1 follow; rows
// K .. 1, the last pass 1 forms, it writes to LDS directly (no checkpoint 0).
// Pass 2 walks the blocks upward: block 0 from those LDS rows, block b > 0 by
// reloading checkpoint b (the next one is already in flight) and
// regenerating the block's K rows into LDS with exactly the operations of
// pass 1 (so every row is bit-identical to the value pass 1 computed); the K
// upward steps consume them.  Scratch traffic per
// point drops from up to 7 doubles per row written and read to at most 7
// doubles per K rows (a ratio chain's state only for the blocks that use
// it), and the upward pass no longer waits on scratch rows.
// Slot layout: checkpoints [b][7][lane] (du.re, du.im, dv.re, dv.im, dw.re,
// dw.im, dy), then the start-correction planes DDV, DDW [row][lane] (complex),
// written only when the reference's start is not converged for v or w.
template <int MINW, int K>
__global__ __launch_bounds__(64, MINW) void k_mie_coated_ckpt(
    const int64_t* __restrict__ order, const int64_t* __restrict__ range, int64_t n_all,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ eps,
    const double* __restrict__ mu, const double* __restrict__ eps2,
    const int64_t* __restrict__ coef_off, double* __restrict__ an, double* __restrict__ bn,
    double* __restrict__ props, int32_t* __restrict__ status, double* __restrict__ scratch,
    int rows_cap, int dn_extra) {
  // rows above n_max in the last block stay below nmx - 1 and every ratio
  // chain's start (CkptPt)
  static_assert(K >= 1 && K <= 12, "checkpoint interval");
  __shared__ double rows[K][7][64];
  const int lane = threadIdx.x;
  const int64_t begin = range[2], end = range[3];
  const int bcap = (rows_cap + K - 1) / K;
  double* slot = scratch + (size_t)blockIdx.x * (size_t)64 * (7 * (size_t)bcap + 4 * rows_cap);
  // groups of 64 sorted points handed out in order (costliest first) by a
  // counter the plan zeroes: a wavefront that finishes early takes the next
  unsigned long long* queue =
      reinterpret_cast<unsigned long long*>(const_cast<int64_t*>(range) + kQueueSlot);
  const int64_t ngroups = (end - begin + 63) / 64;
  double* CK = slot;                                          // [b][7][lane]
  cd* DDV = reinterpret_cast<cd*>(slot + (size_t)7 * bcap * 64);
  cd* DDW = DDV + (size_t)rows_cap * 64;
  // the start-correction planes [DDV | DDW] through a range-checked resource
  const __amdgpu_buffer_rsrc_t rs_dd =
      slot_rsrc(reinterpret_cast<double*>(DDV), (uint32_t)((size_t)rows_cap * 64 * 4 * sizeof(double)));

  for (;;) {
    int64_t grp = 0;
    if (lane == 0) grp = (int64_t)atomicAdd(queue, 1ull);
    grp = __shfl(grp, 0);
    if (grp >= ngroups) break;
    const int64_t base = begin + grp * 64;
    const int64_t idx = base + lane;
    if (idx >= end) continue;
    const int64_t p = order[idx];
    PointDesc d = describe(p, x, y, eps, mu, eps2);
    if (d.cls != 1) continue;
    const double yy = d.size, xx = d.xcore;
    const int nmax = nmax_rule(yy);
    if (nmax > rows_cap) {
      atomicOr(status, 1);
      continue;
    }
    const cd m1 = csqrt_(d.eps), m2 = csqrt_(d.eps2);
    const cd m = cdiv(m2, m1);
    const cd u = m1 * xx, v = m2 * xx, w = m2 * yy;
    const double mxv = fmax(cabs_(m1 * yy), cabs_(w));
    const int nmx =
        nmx_rule(nmax, mxv) + (dn_extra < 0 ? 16 + 10 * (int)ceil(cbrt(mxv)) : dn_extra);
    const double av = cabs_(v), aw = cabs_(w);
    // ratio rows (nsw, s] of each chain; kNoRatio where |z| >= n_max (no ratio
    // rows: the upward form serves every n <= n_max)
    const int nsw_v = (int)av < nmax ? (int)av : kNoRatio;
    const int nsw_w = (int)aw < nmax ? (int)aw : kNoRatio;
    const int nsw_y = (int)yy < nmax ? (int)yy : kNoRatio;
    const cdd ui = crcp_dd(u), vi = crcp_dd(v), wi = crcp_dd(w);
    const ddr yi = rcp_dd(yy);
    const int sv = nsw_v < nmax ? ratio_start(nmax, av) : 0;
    const int sw = nsw_w < nmax ? ratio_start(nmax, aw) : 0;
    const int sy = nsw_y < nmax ? ratio_start(nmax, yy) : 0;
    const int nblk = (nmax + K - 1) / K;
    // ---- pass 1: the four downward chains, checkpointed every K rows
    {
      const int top = max(max(nmx - 1, sv), max(sw, sy));
      cd du = cd{0.0, 0.0}, dv = cd{0.0, 0.0}, dw = cd{0.0, 0.0};
      double dy = 0.0;
      // rows above every ratio chain's start and above the highest
      // checkpoint (row nblk K + 1) advance D_n(u) alone; such rows exist
      // only when top = nmx - 1
      const int lo = max(max(sv, sw), max(sy, nblk * K + 1));
      int k = top;
      for (; k > lo; --k) {
        const cd r = mul_dd((double)(k + 1), ui);
        du = r - crcp_fast(du + r);
      }
      for (; k > K; --k) {
        if (k <= nmx - 1) {   // D_k(u) = (k+1)/u - 1/(D_{k+1} + (k+1)/u), D_nmx = 0
          const cd r = mul_dd((double)(k + 1), ui);
          du = r - crcp_fast(du + r);
        }
        if (k <= sv && k > nsw_v) {
          const cd r = mul_dd((double)k, vi);
          dv = r - crcp_fast(dv + r);
        }
        if (k <= sw && k > nsw_w) {
          const cd r = mul_dd((double)k, wi);
          dw = r - crcp_fast(dw + r);
        }
        if (k <= sy && k > nsw_y) {
          const double r = mul_dd((double)k, yi);
          dy = r - rcp_nr(dy + r);
        }
        const int b = (k - 1) / K - 1;      // state after row bK+K+1
        if ((k - 1) % K == 0 && b >= 1 && b < nblk) {
          // a ratio chain's state only where block b has a row that uses it
          // (a row above the chain's nsw); pass 2 reads it only there
          double* c = CK + (size_t)b * 7 * 64 + lane;
          c[0] = du.re;
          c[64] = du.im;
          if (k - 1 > nsw_v) {
            c[128] = dv.re;
            c[192] = dv.im;
          }
          if (k - 1 > nsw_w) {
            c[256] = dw.re;
            c[320] = dw.im;
          }
          if (k - 1 > nsw_y) c[384] = dy;
        }
      }
      // rows K..1 (pass 2's first block) go to LDS as pass 2 would regenerate
      // them, so block 0 needs no checkpoint and no regeneration
      for (int k = min(top, K); k >= 1; --k) {
        {   // k <= K < nmx - 1, and k <= s for every active chain
          const cd r = mul_dd((double)(k + 1), ui);
          du = r - crcp_fast(du + r);
        }
        rows[k - 1][0][lane] = du.re;
        rows[k - 1][1][lane] = du.im;
        if (k > nsw_v) {
          const cd r = mul_dd((double)k, vi);
          const cd it = crcp_fast(dv + r);
          rows[k - 1][2][lane] = it.re;
          rows[k - 1][3][lane] = it.im;
          dv = r - it;
        }
        if (k > nsw_w) {
          const cd r = mul_dd((double)k, wi);
          const cd it = crcp_fast(dw + r);
          rows[k - 1][4][lane] = it.re;
          rows[k - 1][5][lane] = it.im;
          dw = r - it;
        }
        if (k > nsw_y) {
          const double r = mul_dd((double)k, yi);
          const double it = rcp_nr(dy + r);
          rows[k - 1][6][lane] = it;
          dy = r - it;
        }
      }
    }
    const bool ddv_on = !converged_start(nmx, av), ddw_on = !converged_start(nmx, aw);
    if (ddv_on) {
      dn_chain(vi, nmx, nmax, DDV, lane, false);
      dn_chain(vi, nmx + 16 + 8 * (int)ceil(cbrt(av)), nmax, DDV, lane, true);
    }
    if (ddw_on) {
      dn_chain(wi, nmx, nmax, DDW, lane, false);
      dn_chain(wi, nmx + 16 + 8 * (int)ceil(cbrt(aw)), nmax, DDW, lane, true);
    }
    // ---- pass 2: per block, its K rows in LDS (block 0 left by pass 1, the
    // others regenerated), then K upward steps
    const CkptPt cp{v, w, m, m2, yy, ui, vi, wi, crcp_dd(m), yi,
                    nmax, nsw_v, nsw_w, nsw_y, nblk, ddv_on, ddw_on};
    const int64_t off = coef_off ? coef_off[p] : 0;
    const uint32_t b_ddv = 16u * (uint32_t)lane, b_ddw = 16u * ((uint32_t)rows_cap * 64u + lane);
    Props pr;
    if (__any(ddv_on || ddw_on)) {
      pr = coef_off ? ckpt_pass2<K, true, true>(cp, rows, CK, rs_dd, b_ddv, b_ddw, lane, off, an, bn)
                    : ckpt_pass2<K, true, false>(cp, rows, CK, rs_dd, b_ddv, b_ddw, lane, off, an,
                                                 bn);
    } else {
      pr = coef_off ? ckpt_pass2<K, false, true>(cp, rows, CK, rs_dd, b_ddv, b_ddw, lane, off, an,
                                                 bn)
                    : ckpt_pass2<K, false, false>(cp, rows, CK, rs_dd, b_ddv, b_ddw, lane, off, an,
                                                  bn);
    }
    if (props) pr.finish(yy, p, n_all, props);
  }
}

// ---------------------------------------------------------------------------
// Coated sphere, mie_coeffs.py:183-251 + mie_props.py:28-70 (size = y), literal
// form for absorbing shells (|Im m2| y > 1)
// ---------------------------------------------------------------------------
// Complex Riccati-Bessel upward step with the oscillatory / ratio switch.
struct RBc {
  cd psi_pp, psi_p, chi_pp, chi_p;
  int nsw;
  __device__ __forceinline__ void init(cd z, int nmax) {
    psi_pp = ccos_(z);
    psi_p = csin_(z);
    chi_pp = -psi_p;
    chi_p = psi_pp;
    nsw = min((int)cabs_(z), nmax);
  }
};

__global__ __launch_bounds__(64) void k_mie_coated_literal(
    const int64_t* __restrict__ order, const int64_t* __restrict__ range, int64_t n_all,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ eps,
    const double* __restrict__ mu, const double* __restrict__ eps2,
    const int64_t* __restrict__ coef_off, double* __restrict__ an, double* __restrict__ bn,
    double* __restrict__ props, int32_t* __restrict__ status, double* __restrict__ scratch,
    int rows_cap, int dn_extra) {
  const int lane = threadIdx.x;
  const int64_t begin = range[4], end = range[5];
  // 11 doubles per row: uu, vv, dnw, rv, rw (complex), ry (real)
  double* slot = scratch + (size_t)blockIdx.x * (size_t)rows_cap * 64 * 11;
  const size_t plane = (size_t)rows_cap * 64;
  cd* UU = reinterpret_cast<cd*>(slot);
  cd* VV = reinterpret_cast<cd*>(slot + 2 * plane);
  cd* DW = reinterpret_cast<cd*>(slot + 4 * plane);
  cd* RV = reinterpret_cast<cd*>(slot + 6 * plane);
  cd* RW = reinterpret_cast<cd*>(slot + 8 * plane);
  double* RY = slot + 10 * plane;

  for (int64_t base = begin + (int64_t)blockIdx.x * 64; base < end;
       base += (int64_t)gridDim.x * 64) {
    int64_t idx = base + lane;
    if (idx >= end) continue;
    int64_t p = order[idx];
    PointDesc d = describe(p, x, y, eps, mu, eps2);
    if (d.cls != 2) continue;
    const double yy = d.size, xx = d.xcore;
    const int nmax = nmax_rule(yy);
    if (nmax > rows_cap) {
      atomicOr(status, 1);
      continue;
    }
    cd m1 = csqrt_(d.eps), m2 = csqrt_(d.eps2);
    cd m = cdiv(m2, m1);
    cdd minv = crcp_dd(m);
    cd u = m1 * xx, v = m2 * xx, w = m2 * yy;
    double mxv = fmax(cabs_(m1 * yy), cabs_(w));
    int nmx = nmx_rule(nmax, mxv) + (dn_extra < 0 ? 16 + 10 * (int)ceil(cbrt(mxv)) : dn_extra);
    // three D_n recurrences from D_nmx = 0 sharing one start (mie_coeffs.py:214-218)
    {
      cdd ui = crcp_dd(u), vi = crcp_dd(v), wi = crcp_dd(w);
      cd du = cd{0.0, 0.0}, dv = du, dw = du;
      for (int j = nmx - 1; j >= 1; --j) {
        double jj = (double)(j + 1);
        cd ru = mul_dd(jj, ui), rv = mul_dd(jj, vi), rw = mul_dd(jj, wi);
        du = ru - crcp_fast(du + ru);
        dv = rv - crcp_fast(dv + rv);
        dw = rw - crcp_fast(dw + rw);
        if (j <= nmax) {
          size_t o = (size_t)(j - 1) * 64 + lane;
          UU[o] = m * du - dv;      // mie_coeffs.py:230
          VV[o] = mul_dd(du, minv) - dv;   // mie_coeffs.py:231
          DW[o] = dw;
        }
      }
    }
    RBc bv, bw;
    bv.init(v, nmax);
    bw.init(w, nmax);
    const int nsw_y = min((int)yy, nmax);
    // converged psi ratios where psi decays
    if (bv.nsw < nmax) {
      cdd vi = crcp_dd(v);
      cd dd = cd{0.0, 0.0};
      for (int n = ratio_start(nmax, cabs_(v)); n > bv.nsw; --n) {
        cd r = mul_dd((double)n, vi);
        cd t = dd + r;
        if (n <= nmax) RV[(size_t)(n - 1) * 64 + lane] = t;
        dd = r - crcp_fast(t);
      }
    }
    if (bw.nsw < nmax) {
      cdd wi = crcp_dd(w);
      cd dd = cd{0.0, 0.0};
      for (int n = ratio_start(nmax, cabs_(w)); n > bw.nsw; --n) {
        cd r = mul_dd((double)n, wi);
        cd t = dd + r;
        if (n <= nmax) RW[(size_t)(n - 1) * 64 + lane] = t;
        dd = r - crcp_fast(t);
      }
    }
    const ddr yinv = rcp_dd(yy);
    if (nsw_y < nmax) {
      double dd = 0.0;
      for (int n = ratio_start(nmax, yy); n > nsw_y; --n) {
        double r = mul_dd((double)n, yinv);
        double t = dd + r;
        if (n <= nmax) RY[(size_t)(n - 1) * 64 + lane] = t;
        dd = r - rcp_nr(t);
      }
    }
    double sn, cs;
    sincos(yy, &sn, &cs);
    double py_pp = cs, py_p = sn, cy_pp = -sn, cy_p = cs;
    cdd vinv = crcp_dd(v), winv = crcp_dd(w);
    cdd m2inv = crcp_dd(m2);
    Props pr;
    pr.init();
    const bool store = coef_off != nullptr;
    int64_t off = store ? coef_off[p] : 0;
    // one-row-ahead prefetch of the scratch rows
    cd uu_nx = UU[lane], vv_nx = VV[lane], dw_nx = DW[lane], rv_nx = RV[lane], rw_nx = RW[lane];
    double ry_nx = RY[lane];
    for (int n = 1; n <= nmax; ++n) {
      const cd uu = uu_nx, vv = vv_nx, dnw = dw_nx, rvn = rv_nx, rwn = rw_nx;
      const double ryn = ry_nx;
      if (n < nmax) {
        const size_t o = (size_t)n * 64 + lane;
        uu_nx = UU[o];
        vv_nx = VV[o];
        dw_nx = DW[o];
        rv_nx = RV[o];
        rw_nx = RW[o];
        ry_nx = RY[o];
      }
      double fn = (double)(2 * n - 1);
      // psi/chi at v
      const cd fv_ = mul_dd(fn, vinv), fw_ = mul_dd(fn, winv);
      cd pv = (n <= bv.nsw) ? fv_ * bv.psi_p - bv.psi_pp : cdiv(bv.psi_p, rvn);
      cd chv = fv_ * bv.chi_p - bv.chi_pp;
      bv.psi_pp = bv.psi_p; bv.psi_p = pv; bv.chi_pp = bv.chi_p; bv.chi_p = chv;
      // psi/chi at w
      cd pw = (n <= bw.nsw) ? fw_ * bw.psi_p - bw.psi_pp : cdiv(bw.psi_p, rwn);
      cd chw = fw_ * bw.chi_p - bw.chi_pp;
      bw.psi_pp = bw.psi_p; bw.psi_p = pw; bw.chi_pp = bw.chi_p; bw.chi_p = chw;
      // psi/chi at y (real)
      double f = mul_dd(fn, yinv);
      double py = (n <= nsw_y) ? fma(f, py_p, -py_pp) : py_p * rcp_nr(ryn);
      double chy = fma(f, cy_p, -cy_pp);
      // shell combination, mie_coeffs.py:230-249
      cd fv = cdiv(pv, chv);
      cd fvw = cdiv(fv, pw);
      cd ku1 = uu * fvw;
      cd kv1 = vv * fvw;
      cd pt = pw - chw * fv;
      cd prat = cdiv(cdiv(pw, pv), chv);
      cd ku2 = uu * pt + prat;
      cd kv2 = vv * pt + prat;
      cd dns = cdiv(ku1, ku2) + dnw;
      cd gns = cdiv(kv1, kv2) + dnw;
      double nrat = mul_dd((double)n, yinv);
      cd a1 = mul_dd(dns, m2inv) + nrat;
      cd b1 = m2 * gns + nrat;
      cd gsy = cd{py, -chy}, gs1y = cd{py_p, -cy_p};
      cd a = cdiv(a1 * py - py_p, gsy * a1 - gs1y);
      cd b = cdiv(b1 * py - py_p, gsy * b1 - gs1y);
      pr.add(n, a, b);
      if (store) {
        an[2 * (off + n - 1)] = a.re;
        an[2 * (off + n - 1) + 1] = a.im;
        bn[2 * (off + n - 1)] = b.re;
        bn[2 * (off + n - 1) + 1] = b.im;
      }
      py_pp = py_p; py_p = py; cy_pp = cy_p; cy_p = chy;
    }
    if (props) pr.finish(yy, p, n_all, props);
  }
}

// n_max per point and its exclusive prefix (coefficient offsets), one block.
#define GM_MIE_ARGS                                                                         \
  const int64_t*, const int64_t*, int64_t, const double*, const double*, const double*,    \
      const double*, const double*, const int64_t*, double*, double*, double*, int32_t*,    \
      double*, int, int
template __global__ void k_mie_homog<1>(GM_MIE_ARGS);
template __global__ void k_mie_homog<4>(GM_MIE_ARGS);
template __global__ void k_mie_homog_ckpt<1, 4>(GM_MIE_ARGS);
template __global__ void k_mie_coated_stable<1>(GM_MIE_ARGS);
template __global__ void k_mie_coated_stable<3>(GM_MIE_ARGS);
template __global__ void k_mie_coated_stable<4>(GM_MIE_ARGS);
template __global__ void k_mie_coated_ckpt<1, 4>(GM_MIE_ARGS);
template __global__ void k_mie_coated_ckpt<3, 4>(GM_MIE_ARGS);
template __global__ void k_mie_coated_ckpt<1, 5>(GM_MIE_ARGS);
template __global__ void k_mie_coated_ckpt<1, 8>(GM_MIE_ARGS);
template __global__ void k_mie_homog_coop<4>(GM_MIE_ARGS);
template __global__ void k_mie_homog_coop<16>(GM_MIE_ARGS);
template __global__ void k_mie_homog_coop<32>(GM_MIE_ARGS);
template __global__ void k_mie_homog_coop<64>(GM_MIE_ARGS);
#undef GM_MIE_ARGS

__global__ void k_nmax_of(int64_t n, const double* __restrict__ size, int64_t* __restrict__ cnt) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) cnt[p] = nmax_rule(size[p]);
}

}  // namespace gm
