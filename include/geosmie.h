/*
 * geosmie.h -- C ABI of libgeosmie_hip.so, the gfx950 (MI355X) implementation
 * of GEOSmie's Mie / LUT hot path.
 *
 * Conventions
 *   - Every entry point is extern "C", takes plain pointers and sizes, and
 *     returns an int status: GM_OK (0), < 0 invalid argument, > 0 device or
 *     runtime error.  gm_last_error() returns a message for the calling thread.
 *   - Array arguments are DEVICE pointers (hipMalloc'd memory, e.g. the
 *     data_ptr() of a ROCm torch tensor) unless the name ends in _host.
 *   - Complex values are interleaved (re, im) doubles.
 *   - `stream` is a hipStream_t (NULL = the legacy default stream).  Calls
 *     enqueue work and return without synchronising unless stated.
 *   - A gm_ctx holds grow-on-demand device workspaces; calls on one context
 *     must not overlap (use one context per host thread or stream).
 *
 * The reference (GEOS-ESM/GEOSmie) has no C/FFI boundary on this path: its
 * operator API is Python (pymiecoated and runoptics).  Each entry point below
 * names the reference function it replaces; the Python package geosmie_amd
 * binds these with ctypes and re-exposes the reference's Python names.
 */
#ifndef GEOSMIE_H
#define GEOSMIE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gm_ctx gm_ctx;
typedef void* gm_stream; /* hipStream_t */

enum gm_status {
  GM_OK = 0,
  GM_EINVAL = -1,    /* invalid argument */
  GM_ERUNTIME = 1,   /* HIP runtime error */
  GM_ENOMEM = 2,     /* device allocation failed */
  GM_ECAPACITY = 3   /* a point exceeded the stated size bound */
};

/* gm_mie_batch flags */
enum gm_mie_flags {
  GM_MIE_SORT = 1u,            /* length-bucket the points (recommended) */
  GM_MIE_DN_CONVERGED = 2u     /* start D_n 16 + 10 ceil(|z|^(1/3)) terms above
                                  the reference's nmx (converged also for weakly
                                  absorbing spheres); the default reproduces the
                                  reference's start exactly */
};

/* S1/S2 contraction precision */
enum gm_precision {
  GM_PREC_FP64 = 0,     /* FP64 MFMA */
  GM_PREC_FP32 = 1,     /* FP32 MFMA, FP32 accumulation */
  GM_PREC_BF16X3 = 2    /* split-BF16 (hi*hi + hi*lo + lo*hi), FP32 accumulation */
};

const char* gm_version(void);
const char* gm_last_error(void);

/* Context on HIP device `device` (made current for the calls that use it). */
int gm_ctx_create(int device, gm_ctx** out);
int gm_ctx_destroy(gm_ctx* ctx);
/* Synchronise `stream` and report device-side status flags raised by earlier
 * calls on ctx since the last check: bit 0 = a point exceeded the size bound
 * given to gm_mie_batch (its outputs were not written), bit 1 = a coated point
 * had mu != 1 (the reference raises ValueError, mie_coeffs.py:57-59).
 * Clears the flags. */
int gm_ctx_check(gm_ctx* ctx, gm_stream stream, int32_t* flags_host);

/* Kernel timing (no reference counterpart; instrumentation for the
 * benchmark's roofline figure).  When enabled, calls on ctx record a HIP event
 * pair on their launch stream around each main kernel.  gm_ctx_kernel_time
 * synchronises those events and returns the summed milliseconds and launch
 * count of `kernel` (0 homogeneous Mie, 1 coated Mie, 2 S1/S2 contraction,
 * 3 PSD contraction, 4 GSF expansion, 5 MLP), then forgets them. */
int gm_ctx_set_timing(gm_ctx* ctx, int enable);
int gm_ctx_kernel_time(gm_ctx* ctx, int kernel, double* ms_total, int64_t* launches);

/* Series length n_max = round(2 + s + 4 s^(1/3)) with half-to-even rounding.
 * Replaces the rule at src/pymiecoated/pymiecoated/mie_coeffs.py:99 (and :203
 * with the shell size). Host function. */
int gm_nmax(double size);

/* Coefficient offsets for a batch: off[0] = 0, off[i+1] = off[i] + n_max(size[i]).
 * size = x (homogeneous) or y (coated).  off has n+1 entries (device). */
int gm_coef_offsets(gm_ctx* ctx, int64_t n, const double* size, int64_t* off, gm_stream stream);

/* Mie coefficients a_n, b_n and efficiencies for n points, FP64.
 *
 * Replaces mie_coeffs (src/pymiecoated/pymiecoated/mie_coeffs.py:36-73) with
 * single_mie_coeff[_numba] (:83-180) and coated_mie_coeff (:183-251), plus
 * mie_props_raw (src/pymiecoated/pymiecoated/mie_props.py:28-70), as looped
 * over sizes by MultipleMie.calculateS12SizeRange
 * (src/pymiecoated/pymiecoated/mie_coated.py:61-89).
 *
 *   x     [n]    size parameter (core size parameter when y is given)
 *   y     [n]    shell size parameter, or NULL for homogeneous spheres
 *   eps   [2n]   relative permittivity (core), complex
 *   mu    [2n]   relative permeability, complex, or NULL for mu = 1
 *   eps2  [2n]   shell permittivity (required iff y != NULL)
 *   size_max     upper bound on max(x) (max(y) when coated) over the batch;
 *                sizes workspaces.  Points above it raise status bit 0.
 *   coef_off [n+1]  offsets from gm_coef_offsets, or NULL to skip a_n/b_n
 *   an, bn   complex [coef_off[n]]  outputs (ignored when coef_off is NULL)
 *   props [6*n]  qext, qsca, qabs, qb, asy, qratio planes (props[k*n + i]),
 *                or NULL.  The size in the efficiencies is y for coated
 *                points, x otherwise (mie_coated.py:196).
 *   nmax_out [n] int32 n_max per point, or NULL
 * Dispatch per point as mie_coeffs.py:61-71 (x == y or eps == eps2 ->
 * homogeneous with eps at y; x == 0 -> homogeneous with eps2). */
int gm_mie_batch(gm_ctx* ctx, int64_t n, const double* x, const double* y, const double* eps,
                 const double* mu, const double* eps2, double size_max, const int64_t* coef_off,
                 double* an, double* bn, double* props, int32_t* nmax_out, uint32_t flags,
                 gm_stream stream);

/* Efficiencies of given coefficient sets: qext, qsca, qabs, qb, asy, qratio
 * planes in props[6*n].  Replaces mie_props / mie_props_raw
 * (src/pymiecoated/pymiecoated/mie_props.py:28-75).  size[n] is the size
 * parameter used in the normalisation (y for coated particles). */
int gm_mie_props(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const double* an,
                 const double* bn, const double* size, double* props, gm_stream stream);

/* Premultiplied angle functions pi'_n, tau'_n = (2n+1)/(n(n+1)) * pi_n, tau_n
 * for n = 1..nmax_cap at n_ang cosines, layout [n-1][angle].  Replaces
 * mie_p / mie_t / mie_ptnumba (src/pymiecoated/pymiecoated/mie_props.py:166-231)
 * and MultipleMie.preCalculatePT (mie_coated.py:160-179): one table serves
 * every size (the reference's per-nmax tables are its prefixes). */
int gm_angle_tables(gm_ctx* ctx, int n_ang, const double* cos_theta, int nmax_cap, double* pi_t,
                    double* tau_t, gm_stream stream);

/* S1, S2 (complex, [n][n_ang]) of n coefficient sets against an angle table.
 * Replaces mie_S12_backend_pt / numba_dot (mie_props.py:133-149) as looped by
 * runS12Loop (mie_coated.py:31-39).  nmax[n] int32 per set; the table must
 * have nmax_cap >= max(nmax) rows.  precision: gm_precision. */
int gm_mie_s12(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
               const double* an, const double* bn, int n_ang, const double* pi_t,
               const double* tau_t, int nmax_cap, double* s1, double* s2, int precision,
               gm_stream stream);

/* Fused S1/S2 -> Mueller elements: mueller[n][4][n_ang] = s11, s12, s33, s34
 * per point (s22 = s11, s44 = s33).  Replaces the S12 loop plus
 * calculateScatVals (src/geosmie/dointegration.py:1044-1050, 1223-1235). */
int gm_s12_mueller(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
                   const double* an, const double* bn, int n_ang, const double* pi_t,
                   const double* tau_t, int nmax_cap, double* mueller, int precision,
                   gm_stream stream);

/* Weighted sums over the size grid: out[k][c] = sum_p w[k][p] * v[p][c] for
 * n_psd weight rows w [n_psd][n_pts] and per-size columns v [n_pts][n_cols].
 * The integration of every phase-matrix element in integratePSD
 * (src/geosmie/dointegration.py:1164-1166; v = gm_s12_mueller output viewed
 * as [n_pts][4*n_ang]) and the LUT sweep's batched PSDs.  FP64; MFMA tiles
 * when n_psd > 8, split-K otherwise; fixed reduction order. */
int gm_psd_contract(gm_ctx* ctx, int n_psd, int64_t n_pts, int n_cols, const double* w,
                    const double* v, double* out, gm_stream stream);

/* integratePSD scalars (src/geosmie/dointegration.py:1064-1209, mixing
 * 'dumix', including its Qsca weighting of g, csca, cext from the aliased
 * thisweight, :1111-1112 and :1154-1157, and the qb back-out, :1133-1169).
 *   xx [n_pts] size parameters; props [6][n_pts] from gm_mie_batch;
 *   psd [n_modes][n_pts] number weights; fracs, reff0 [n_modes];
 *   lam wavelength (same unit as the radii); rhop0, rhop densities.
 * out [16]: num, area, volume, mass, rEff, rMass, qext, qsca, qabs, qb, g,
 *   csca, cext, bsca, bext, bbck (mode-summed with fracs, :1199-1200). */
int gm_integrate_psd(gm_ctx* ctx, int64_t n_pts, int n_modes, const double* xx, const double* props,
                     const double* psd, const double* fracs, const double* reff0, double lam,
                     double rhop0, double rhop, double* out, gm_stream stream);

/* Phase-matrix normalisation (dointegration.py:978-988) in place on
 * pmat [n_entries][4][n_ang] (s11, s12, s33, s34): P11 -> 2 P11 / trapz(P11
 * sin(theta), theta), other elements scaled by the same factor; pback
 * [n_entries][6] (or NULL) = values at the last angle in the order p11, p12,
 * p33, p34, p22, p44.  theta_rad [n_ang] device. */
int gm_pmatrix_normalize(gm_ctx* ctx, int64_t n_entries, int n_ang, const double* theta_rad,
                         double* pmat, double* pback, gm_stream stream);

/* Generalised-spherical-function expansion of n_entries phase matrices.
 * Replaces spher_expan.x (src/gsf/spher_expan.f) as driven per (bin,
 * wavelength, rh) by convertncdf.py:170-189, 369-394.
 *   ang_deg_host [n_ang] increasing scattering angles in degrees (host);
 *   F [n_entries][6][n_ang] elements in table order p11, p12, p33, p34, p22, p44;
 *   ng  number of Gauss nodes = number of coefficients (NSPHER = 129 in the
 *       reference, params.h:14);
 *   quantize  nonzero: round to 10 decimals like the F17.10 text output;
 *   out [n_entries][6][ng] normalised coefficients, same element order
 *       (alpha1, beta1, alpha3, beta2, alpha2, alpha4), i.e. pmom[p][m]. */
int gm_gsf_expand(gm_ctx* ctx, int64_t n_entries, int n_ang, const double* ang_deg_host,
                  const double* F, int ng, int quantize, double* out, gm_stream stream);

/* out[s][i] = sum_l coef[s][l] P_l(x[i]) for n_series series of L terms
 * (hbleg, src/gsf/convertncdf.py:32-43). */
int gm_legendre_eval(gm_ctx* ctx, int64_t n_x, const double* x, int n_series, int L,
                     const double* coef, double* out, gm_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* GEOSMIE_H */
