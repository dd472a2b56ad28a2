// Shared declarations of the libgeosmie_hip kernels (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gm {

// Counting-sort buckets for the recurrence-length sort: cost >> kBucketShift.
constexpr int kBucketShift = 2;
constexpr int kBuckets = 1024;
// point classes: 0 homogeneous, 1 coated (real / weakly absorbing shell),
// 2 coated (absorbing shell)
constexpr int kClasses = 3;
// points per thread in the plan kernels
constexpr int kPlanPer = 16;

__global__ void k_zero_i64(int64_t* p, int n);
__global__ void k_zero_i32(int32_t* p, int n);
__global__ void k_plan_keys(int64_t n, const double* x, const double* y, const double* eps,
                            const double* mu, const double* eps2, int32_t* key,
                            int32_t* nmax_out, int64_t* hist, int32_t* status, int per);
__global__ void k_plan_scan(const int64_t* hist, int64_t* cursor, int64_t* range, int small_b,
                            int64_t* ab);
__global__ void k_plan_scatter(int64_t n, const int32_t* key, int64_t* cursor, int64_t* order,
                               int per);
__global__ void k_plan_identity(int64_t n, int64_t* order, int64_t* range, int64_t* ab);
// homogeneous points of cost below this (bucketed) run 4 lanes per point: then
// 3 n_max + 16 <= cost < 160, so n_max < 48 and at most 12 rows per lane
constexpr int kSmallCost = 160;
// Plan workspace after the histogram and cursors: range[2 kClasses], the
// homogeneous split ab[4], then one work-queue counter; kQueueSlot is the
// counter's offset from range.
constexpr int kQueueSlot = 2 * kClasses + 4;
constexpr int kSmallLcap = 12;
template <int MINW>
__global__ void k_mie_homog(const int64_t* order, const int64_t* range, int64_t n_all,
                            const double* x, const double* y, const double* eps, const double* mu,
                            const double* eps2, const int64_t* coef_off, double* an, double* bn,
                            double* props, int32_t* status, double* scratch, int rows_cap,
                            int dn_extra);
template <int MINW, int K>
__global__ void k_mie_homog_ckpt(const int64_t* order, const int64_t* range, int64_t n_all,
                                 const double* x, const double* y, const double* eps,
                                 const double* mu, const double* eps2, const int64_t* coef_off,
                                 double* an, double* bn, double* props, int32_t* status,
                                 double* scratch, int rows_cap, int dn_extra);
template <int MINW>
__global__ void k_mie_coated_stable(const int64_t* order, const int64_t* range, int64_t n_all,
                                    const double* x, const double* y, const double* eps,
                                    const double* mu, const double* eps2, const int64_t* coef_off,
                                    double* an, double* bn, double* props, int32_t* status,
                                    double* scratch, int rows_cap, int dn_extra);
template <int P>
__global__ void k_mie_homog_coop(const int64_t* order, const int64_t* range, int64_t n_all,
                                 const double* x, const double* y, const double* eps,
                                 const double* mu, const double* eps2, const int64_t* coef_off,
                                 double* an, double* bn, double* props, int32_t* status,
                                 double* scratch, int lcap, int dn_extra);
template <int MINW, int K>
__global__ void k_mie_coated_ckpt(const int64_t* order, const int64_t* range, int64_t n_all,
                                  const double* x, const double* y, const double* eps,
                                  const double* mu, const double* eps2, const int64_t* coef_off,
                                  double* an, double* bn, double* props, int32_t* status,
                                  double* scratch, int rows_cap, int dn_extra);
__global__ void k_mie_coated_literal(const int64_t* order, const int64_t* range, int64_t n_all,
                                     const double* x, const double* y, const double* eps,
                                     const double* mu, const double* eps2, const int64_t* coef_off,
                                     double* an, double* bn, double* props, int32_t* status,
                                     double* scratch, int rows_cap, int dn_extra);
__global__ void k_nmax_of(int64_t n, const double* size, int64_t* cnt);

// Doubles of scratch per row per lane.
constexpr int kHomogRowDoubles = 3;
constexpr int kCoatedRowDoubles = 11;
// Doubles per lane of a checkpointed coated slot: 7 per K-row block + the
// two complex start-correction planes.
inline int64_t coated_ckpt_lane_doubles(int rows_cap, int k) {
  return 7 * (int64_t)((rows_cap + k - 1) / k) + 4 * (int64_t)rows_cap;
}
// Doubles per lane of a checkpointed homogeneous slot: 3 per K-row block.
inline int64_t homog_ckpt_lane_doubles(int rows_cap, int k) {
  return 3 * (int64_t)((rows_cap + k - 1) / k);
}

}  // namespace gm
