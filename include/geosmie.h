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

/* integratePSD quirks (GM_TUNE_PSD_QUIRKS bits) */
enum gm_psd_quirks {
  GM_PSD_QUIRK_ALIAS = 1u   /* csca and cext weighted by area x Qsca, as the reference's aliased
                               thisweight leaves them after key 'g' (dointegration.py:1111-1112,
                               1154-1157); clear for area weighting like qext, qsca, qabs.  (The
                               qb back-out, :1133-1169, equals the area-weighted average of Qb
                               in exact arithmetic and needs no switch.) */
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

/* Launch-shape parameters of a context (no reference counterpart), set by
 * an autotuner (geosmie_amd/autotune.py) from timings of the caller's own
 * workload.  Results do not depend on them (fixed reduction orders), except
 * at rounding level for GM_TUNE_PSD_CHUNK, which regroups the split-K partial
 * sums of the few-PSD contraction, and GM_TUNE_MIE_WAVES, whose register
 * budgets let the compiler fuse a different set of multiply-adds. */
enum gm_tune_param {
  GM_TUNE_MIE_GRID_MULT = 0,  /* persistent Mie grid = this x resident wavefronts (1..16; 1) */
  GM_TUNE_S12_WAVES = 1,      /* wavefronts of 64 angles per S1/S2 block (1, 2, 4; 4) */
  GM_TUNE_PSD_CHUNK = 2,      /* points per split-K chunk, few-PSD contraction (16..2^20; 128) */
  GM_TUNE_MLP_WEIGHT_TERMS = 3, /* precision, not shape: BF16 terms per MLP weight on the matrix
                                   cores, 1 = BF16 weights (default), 2 = hi + lo BF16 pair */
  GM_TUNE_S12_POINTS = 4,       /* points per wavefront of the FP64 S1/S2 kernel (8, 16; 8) */
  GM_TUNE_MIE_WAVES = 5         /* Mie recurrence kernels' register budget: 0 = unconstrained
                                   (default), 3 or 4 = at least that many resident wavefronts
                                   per SIMD (fewer registers, a few spills) */,
  GM_TUNE_S12_LAYOUT = 6        /* FP64 S1/S2 blocks: 0 = the block's wavefronts take different
                                   angles of the same points (default), 1 = different points
                                   of the same 64 angles (angle-table rows shared in cache),
                                   2 = as 1 with the angle tables staged through LDS */,
  GM_TUNE_MIE_CKPT = 7          /* coated kernel (real / weakly absorbing shells): the downward
                                   chains' states kept every 4, 5 or 8 rows and each block of rows
                                   regenerated in LDS before use (4, the default), or 0 = every
                                   row kept in scratch; results agree to rounding.  Non-zero also
                                   selects the checkpointed one-point-per-lane homogeneous kernel
                                   (every 4 rows, unless GM_TUNE_MIE_WAVES forces a register
                                   budget; bit-identical results) */,
  GM_TUNE_MIE_LANES = 8         /* homogeneous kernel: lanes per point.  1 = one point per lane
                                   (large batches); 16, 32, 64 = each point's rows split over that
                                   many lanes, its serial recurrences joined by parallel scans
                                   (batches too small to fill the GPU); 0 (default) = chosen from
                                   the batch size and series length.  Results agree to rounding */,
  GM_TUNE_PSD_QUIRKS = 9,       /* a behaviour setting, not tuned for speed: the integratePSD
                                   quirks gm_integrate_psd[_rows] reproduce (GM_PSD_QUIRK_* bits;
                                   default GM_PSD_QUIRK_ALIAS) */
  GM_TUNE_S12_ROW_TILES = 10,   /* gm_s12_contract_rows: 8-point tiles per work unit (1, 2, 4): at
                                   FP64 one after the other, at FP32 / split-BF16 multiplied
                                   together, sharing each angle-table fetch; 0 (default) = by
                                   batch size.  Results agree to rounding */
  GM_TUNE_S12_SHARE = 11        /* gm_s12_contract_rows at split-BF16 with one tile per work
                                   unit: 1 (default) = four work units per block share each
                                   K-step's angle-table tile through LDS; 0 = each wavefront
                                   fetches its own.  Identical results */
};
int gm_ctx_tune(gm_ctx* ctx, int param, int value);
int gm_ctx_get_tune(gm_ctx* ctx, int param, int* value);

/* Kernel timing (no reference counterpart; instrumentation for the
 * benchmark's roofline figure).  When enabled, calls on ctx record a HIP event
 * pair on their launch stream around each main kernel.  gm_ctx_kernel_time
 * synchronises those events and returns the summed milliseconds and launch
 * count of `kernel` (0 homogeneous Mie, 1 coated Mie, 2 S1/S2 contraction,
 * 3 PSD contraction, 4 GSF expansion, 5 MLP, 6 coated Mie with absorbing
 * shells, 7 LUT bulk reduction, 8 S1/S2 operand conversion), then forgets
 * them. */
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

/* gm_mie_batch without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * Same arguments, outputs and kernels as gm_mie_batch, plus `status`: a
 * caller-owned device int32 that the call clears and then ORs the status bits
 * of gm_ctx_check into (bit 0: a point exceeded size_max; bit 1: a coated
 * point with mu != 1).  The plan buffers and the recurrence slots live in one
 * static device array of 128 MiB per GPU, held until the process ends; the
 * call itself allocates nothing (ctx must exist: gm_ctx_create, which may
 * allocate, takes the array's address on its device).  Batches whose plan
 * (12 bytes per point plus 49 KiB) and one recurrence slot (n_max(size_max)
 * x 64 lanes x 24 bytes, or 88 bytes when coated) exceed that array are
 * rejected with GM_EINVAL; smaller slot budgets only shorten the persistent
 * grid.  Every call on a GPU shares that GPU's array, which keeps the last
 * call's contents: calls that overlap on one GPU (other streams, threads or
 * contexts) race and must be serialised by the caller. */
int gm_mie_batch_noalloc(gm_ctx* ctx, int64_t n, const double* x, const double* y,
                         const double* eps, const double* mu, const double* eps2,
                         double size_max, const int64_t* coef_off, double* an, double* bn,
                         double* props, int32_t* nmax_out, uint32_t flags, int32_t* status,
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

/* P11 only: p11[n][n_ang] = (|S1|^2 + |S2|^2) / 2 (s11 of calculateScatVals,
 * dointegration.py:1045), for the LUT sweep, which needs no other element. */
int gm_s12_p11(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
               const double* an, const double* bn, int n_ang, const double* pi_t,
               const double* tau_t, int nmax_cap, double* p11, int precision, gm_stream stream);

/* Weighted sums over the size grid: out[k][c] = sum_p w[k][p] * v[p][c] for
 * n_psd weight rows w [n_psd][n_pts] and per-size columns v [n_pts][n_cols].
 * The integration of every phase-matrix element in integratePSD
 * (src/geosmie/dointegration.py:1164-1166; v = gm_s12_mueller output viewed
 * as [n_pts][4*n_ang]) and the LUT sweep's batched PSDs.  FP64; MFMA tiles
 * when n_psd > 8, split-K otherwise; fixed reduction order. */
int gm_psd_contract(gm_ctx* ctx, int n_psd, int64_t n_pts, int n_cols, const double* w,
                    const double* v, double* out, gm_stream stream);

/* Rows over their own point ranges: out[r][c] = sum over p in
 * [start_host[r], start_host[r] + count_host[r]) of w[r][p] v[p][c], for
 * n_rows <= 64 rows (one runoptics slab per wavelength: integratePSD's
 * phase-matrix sums, dointegration.py:1164-1166).  w [n_rows][n_pts] and v
 * [n_pts][n_cols] device; the ranges are host arrays.  Split-K over chunks of
 * GM_TUNE_PSD_CHUNK points, chunks of a row added in order. */
int gm_psd_contract_rows(gm_ctx* ctx, int n_rows, const int64_t* start_host,
                         const int64_t* count_host, int64_t n_pts, int n_cols, const double* w,
                         const double* v, double* out, gm_stream stream);

/* gm_s12_mueller followed by gm_psd_contract_rows, fused: out[r][e][j]
 * = sum over p in [start_host[r], start_host[r] + count_host[r]) of w[r][p]
 * s_e(p, j), e = s11, s12, s33, s34, without materialising the Mueller planes
 * (calculateScatVals then integratePSD's phase-matrix sums,
 * src/geosmie/dointegration.py:1044-1050, 1164-1166).  Arguments as
 * gm_s12_mueller (coefficients of n points) and gm_psd_contract_rows (n_rows
 * <= 64, ranges on the host, w [n_rows][n] device); out [n_rows][4][n_ang].
 * precision: GM_PREC_FP64 reads an, bn; GM_PREC_FP32 and GM_PREC_BF16X3 read
 * coef_lp, the same coefficients converted by gm_s12_split_coef (an, bn may
 * then be NULL); the Mueller elements and the sums are FP64 at every
 * precision.  Consecutive rows with the same range share one S1/S2 pass (up
 * to 4 rows each), so it suits rows that overlap little, as runoptics'
 * per-rh rows do; many rows over the same points are cheaper through
 * gm_s12_mueller and gm_psd_contract.  Fixed summation order. */
int gm_s12_contract_rows(gm_ctx* ctx, int64_t n, const int64_t* coef_off, const int32_t* nmax,
                         const double* an, const double* bn, int n_ang, const double* pi_t,
                         const double* tau_t, int nmax_cap, int n_rows, const int64_t* start_host,
                         const int64_t* count_host, const double* w, double* out, int precision,
                         const void* coef_lp, gm_stream stream);

/* The coefficients of n_coef (a_n, b_n) pairs (an, bn: 2 n_coef doubles,
 * indexed as gm_mie_batch writes them) in the form the reduced-precision
 * S1/S2 passes multiply: U = a + b, V = a - b, formed in FP64, then as FP32
 * (precision GM_PREC_FP32) or as BF16 hi + lo pairs (GM_PREC_BF16X3); 16
 * bytes per pair into coef_lp (device, 16 n_coef bytes).  Converting once
 * per call rather than in every angle wavefront is what this saves
 * (mie_props.py:133-149 forms the same sums per angle). */
int gm_s12_split_coef(gm_ctx* ctx, int64_t n_coef, const double* an, const double* bn,
                      int precision, void* coef_lp, gm_stream stream);

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

/* integratePSD for n_rows independent single-mode rows at once (one
 * runoptics slab per wavelength, fracs = 1): psd [n_rows][n_pts] number
 * weights, each row nonzero on its own points; lam_pts [n_pts] the wavelength
 * of each point's row; reff0 [n_rows]; rhop_rows [n_rows] device, each row's
 * wet density (runoptics: one row per rh, dointegration.py:867-868), or NULL
 * for rhop on every row; out [n_rows][16] as gm_integrate_psd. */
int gm_integrate_psd_rows(gm_ctx* ctx, int64_t n_rows, int64_t n_pts, const double* xx,
                          const double* lam_pts, const double* props, const double* psd,
                          const double* reff0, double rhop0, double rhop,
                          const double* rhop_rows, double* out, gm_stream stream);

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
 *   quantize  bit 0: round to 10 decimals like the F17.10 text output;
 *             bit 1 (GM_GSF_RAW): the coefficients before the 1/alpha1(0)
 *             normalisation (the AL1..BET2 the program's MATR reads);
 *   out [n_entries][6][ng] normalised coefficients, same element order
 *       (alpha1, beta1, alpha3, beta2, alpha2, alpha4), i.e. pmom[p][m]. */
int gm_gsf_expand(gm_ctx* ctx, int64_t n_entries, int n_ang, const double* ang_deg_host,
                  const double* F, int ng, int quantize, double* out, gm_stream stream);

/* The fitted phase matrix back from expansion coefficients: MATR
 * (src/gsf/spher_expan.f:419-517), which the program evaluates at the input
 * angles for '<file>.expan_matr' (:85-91) and for its fit error (ERREVAL,
 * :635-687, here left to the caller).  coef [n_entries][6][l1max] in
 * gm_gsf_expand's order; theta_rad [n_ang] (device); out [n_entries][6][n_ang]
 * in the table order p11, p12, p33, p34, p22, p44.  n_entries <= 65535. */
int gm_gsf_matr(gm_ctx* ctx, int64_t n_entries, int l1max, const double* coef, int n_ang,
                const double* theta_rad, double* out, gm_stream stream);

/* Band means of table spectra (bandaverage.doAverage, src/geosmie/
 * bandaverage.py:18-50): out[s][b] = mean over the n_sub (<= 128) sample
 * wavenumbers samples[b][0..n_sub) of numpy.interp(sample, wavenumber,
 * y[s]) (the rule scipy's interp1d applies to 1-D float64 data), summed in
 * numpy's pairwise order; bit-identical to the reference.  wavenumber
 * [n_lam] ascending, y [n_series][n_lam] in that order, samples already
 * clipped to the wavenumber range (bandaverage.py:38-41); all device. */
int gm_band_average(gm_ctx* ctx, int64_t n_series, int n_lam, const double* wavenumber,
                    const double* y, int n_band, int n_sub, const double* samples, double* out,
                    gm_stream stream);

/* out[s][i] = sum_l coef[s][l] P_l(x[i]) for n_series series of L terms
 * (hbleg, src/gsf/convertncdf.py:32-43). */
int gm_legendre_eval(gm_ctx* ctx, int64_t n_x, const double* x, int n_series, int L,
                     const double* coef, double* out, gm_stream stream);

/* ---- LUT sweep --------------------------------------------------------------
 * A slab is one (m_dry, wavelength, growth factor) combination over a dry
 * radius grid: one iteration of runoptics' (lambda, rh) loop
 * (src/geosmie/dointegration.py:787-889).  Points are slab-major. */

/* Points of n_slab slabs over r_dry[n_r]: x[s*n_r + p] = 2 pi gf[s] r_dry[p] / lam[s]
 * and eps = m^2 with the volume mixing of dointegration.py:497-507,
 * m = m_w + (m_dry[s] - m_w) (1/gf[s])^3.  m_dry complex [2*n_slab]; lam,
 * gf [n_slab]; x [n_slab*n_r]; eps complex [2*n_slab*n_r].  Replaces the
 * per-(lambda, rh) setup of runoptics (dointegration.py:813-832). */
int gm_lut_points(gm_ctx* ctx, int64_t n_slab, int n_r, const double* r_dry, const double* lam,
                  const double* m_dry, const double* gf, double mw_re, double mw_im, double* x,
                  double* eps, gm_stream stream);

/* Truncated lognormal weight rows, as runoptics' calculatePSD builds one
 * mode (getLogNormPSD, src/geosmie/particleparams.py:113-127, times the
 * widths dr, normalised to sum 1: src/geosmie/dointegration.py:569-577):
 * w[k][col0[k] + p] = n(r_p) dr_p / sum_p n(r_p) dr_p for p < n_r, with
 * n(r) = exp(-ln^2(r / rmode) / (2 ln^2 sigma)) / (r sqrt(2 pi) ln sigma),
 * zero where r_p >= rmax or r_p <= rmin.  par [n_rows][4] = (rmode,
 * ln sigma, rmin, rmax) in the units of r (size parameters: radii times
 * 2 pi / lambda); col0 [n_rows] column offsets.  All arrays on the device;
 * w has n_rows rows of ldw >= n_r entries, of which each row's n_r from
 * col0 are written (a row whose offset does not fit is left as it is).
 * One pass per table wavelength replaces the host PSDs of every rh. */
int gm_lognorm_rows(gm_ctx* ctx, int64_t n_rows, int n_r, const double* r, const double* dr,
                    const double* par, const int64_t* col0, double* w, int64_t ldw,
                    gm_stream stream);

/* Lognormal number weights on a radius grid, one row per distribution:
 * w[k][p] = n(r_p; rmode[k], sigma[k]) dr[p] / sum_p(...) with the density of
 * src/geosmie/particleparams.py:113-127 and the bin widths dr (getDR,
 * dointegration.py:93-101), normalised as calculatePSD does (:569-577). */
int gm_lut_lognormal(gm_ctx* ctx, int64_t n_psd, int n_r, const double* r, const double* dr,
                     const double* rmode, const double* sigma, double* w, gm_stream stream);

/* Bulk optics of n_psd size distributions per slab, float32 rows
 * out[s][k][3 + n_mom] = Qext, SSA, g, alpha1_0 .. alpha1_{n_mom-1}:
 * integratePSD's area-weighted Qext/Qsca and Qsca-weighted g
 * (dointegration.py:1102-1190, 961) and the first n_mom Legendre moments of
 * P11 as spher_expan.f forms alpha1 with ng Gauss nodes (SPHER_EXPAN
 * :313-347), normalised by alpha1_0 (:95-103).
 *   w: number weights [n_psd][n_r] shared by all slabs (w_per_slab = 0) or
 *      [n_slab][n_psd][n_r] (w_per_slab = 1);
 *   x [n_slab*n_r], lam [n_slab], props [6][n_slab*n_r] (gm_mie_batch),
 *   p11 [n_slab*n_r][n_ang] (gm_s12_p11); ang_deg_host [n_ang] increasing;
 *   1 <= n_mom <= min(ng, 64). */
int gm_lut_bulk(gm_ctx* ctx, int64_t n_slab, int n_r, int n_ang, int n_psd, const double* w,
                int w_per_slab, const double* x, const double* lam, const double* props,
                const double* p11, const double* ang_deg_host, int ng, int n_mom, float* out,
                gm_stream stream);

/* ---- MLP surrogate (no reference counterpart) --------------------------------
 * A fully connected network dims[0] -> dims[1] -> ... -> dims[n_layers] with
 * SiLU after every layer but the last.  params: one FP32 vector, per layer
 * W [out][in] row-major then b [out] (gm_mlp_param_count entries).  The
 * model's weights are BF16 (every product uses the round-to-nearest-even BF16
 * of the FP32 master value); activations are FP32, fed to the matrix cores as
 * a two-term BF16 split; accumulation is FP32.  dims is a host array of
 * n_layers + 1 widths (each 1..4096, n_layers <= 16).  x [n][dims[0]],
 * y / target [n][dims[n_layers]], row-major FP32, device. */
int64_t gm_mlp_param_count(int n_layers, const int* dims);

/* y = net(x). */
int gm_mlp_forward(gm_ctx* ctx, int n_layers, const int* dims, int64_t n, const float* params,
                   const float* x, float* y, gm_stream stream);

/* Loss sum_j out_weight[j] * mean_b (y[b][j] - target[b][j])^2 into loss[0]
 * (device float) and its gradient with respect to params into grads. */
int gm_mlp_grad(gm_ctx* ctx, int n_layers, const int* dims, int64_t n, const float* params,
                const float* x, const float* target, const float* out_weight, float* grads,
                float* loss, gm_stream stream);

/* One Adam step (bias-corrected, step >= 1) on params with moments m, v. */
int gm_mlp_adam(gm_ctx* ctx, int64_t n_params, float* params, const float* grads, float* m,
                float* v, float lr, float beta1, float beta2, float eps, int64_t step,
                gm_stream stream);

/* Surrogate queries (no reference counterpart).  q [n][6] = (ln r_eff,
 * sigma_g, m_r, ln m_i, lambda, f_w): wet effective radius [um] of a lognormal
 * number distribution (r_eff = gf r_mode exp(2.5 ln^2 sigma_g)), its geometric
 * width, the dry refractive index, the wavelength [um] and the water volume
 * fraction f_w = 1 - gf^-3 of the humidity growth (dointegration.py:497-507).
 * x [n][6]: network features (ln(2 pi r_eff / lambda), sigma_g, Re m_mix,
 * ln Im m_mix, ln(gf / lambda), f_w) with the volume-mixed index
 * m_mix = m_w + (m - m_w)(1 - f_w) (real water index m_water_re), standardised
 * with feat_mean/feat_std.  flags [n]: 0 = served by the network,
 * 1 = outside the box [box_lo, box_hi] of the gate coordinates (ln r_mode,
 * sigma_g, m_r, ln m_i, ln lambda, f_w), 2 = inside but in a cell of the
 * uniform n_cells_host[6] grid whose cell_ok byte is 0.  cell_ok holds
 * n_cell_ok bytes, which must equal the product of n_cells_host (checked).
 * All arrays but n_cells_host are device arrays. */
int gm_surrogate_prep(gm_ctx* ctx, int64_t n, const float* q, const float* feat_mean,
                      const float* feat_std, const float* box_lo, const float* box_hi,
                      const int* n_cells_host, const uint8_t* cell_ok, int64_t n_cell_ok,
                      double m_water_re, float* x, int32_t* flags, gm_stream stream);

/* Network outputs y [n][n_out] (standardised) -> table rows [n][n_out]
 * (Qext, SSA, g, alpha1_0..): Qext = B exp(t0), SSA = min(exp(t1), 1),
 * g = exp(t2), alpha1_l = (2l + 1) t_{3+l}, with t_j = y_j tgt_std[j] + tgt_mean[j].
 * B is the bulk Rayleigh extinction of query q[i] (device [n][6], the
 * gm_surrogate_prep queries) saturated at 2:
 *   Q_R = 8/3 |K|^2 x_eff^4 exp(6 ln^2 sigma_g) + 4 Im(K) x_eff,  B = Q_R / (1 + Q_R / 2),
 * K = (m^2 - 1)/(m^2 + 2) at the index m_water_re + (m_r - m_water_re)(1 - f_w)
 * + i m_i (1 - f_w); q = NULL gives B = 1 (models trained without it). */
int gm_surrogate_post(gm_ctx* ctx, int64_t n, int n_out, const float* q, double m_water_re,
                      const float* y, const float* tgt_mean, const float* tgt_std, float* rows,
                      gm_stream stream);

/* Inverse of gm_surrogate_post: table rows [n][n_out] -> standardised
 * training targets y [n][n_out] (t0 = ln Qext - ln B, t1 = ln SSA, t2 = ln max(g, 1e-8),
 * t_{3+l} = alpha1_l / (2l + 1); y_j = (t_j - tgt_mean[j]) / tgt_std[j]). */
int gm_surrogate_targets(gm_ctx* ctx, int64_t n, int n_out, const float* q, double m_water_re,
                         const float* rows, const float* tgt_mean, const float* tgt_std, float* y,
                         gm_stream stream);

/* ---- GRASP kernel mode (runoptics 'mode': 'kernel') ------------------- */

/* Shape averaging of kernel tables: out[j] = sum_r w[r] data[r][j] over
 * n_shape shapes (r ascending), data [n_shape][n_elem], w [n_shape] device
 * weights.  Replaces integrateShapes (src/geosmie/dointegration.py:74-91),
 * whose weights are the shape fractions normalised to sum 1. */
int gm_shape_average(gm_ctx* ctx, int n_shape, int64_t n_elem, const double* w,
                     const double* data, double* out, gm_stream stream);

/* Bilinear interpolation in refractive index of a table [n_mr][n_mi][n_elem]
 * at n_q queries: idx[2q], idx[2q+1] = the bracketing (m_r, m_i) indices,
 * w[4q..4q+3] = the weights of the corners (i, j), (i+1, j), (i+1, j+1),
 * (i, j+1); out[q][k] = ((w0 f(i,j) + w1 f(i+1,j)) + w3 f(i,j+1)) + w2 f(i+1,j+1),
 * each product rounded, as numpy evaluates bilinear_interpolation
 * (dointegration.py:29-38).  Indices and weights come from get_interpolated's
 * rules (:40-72: bisect, clamped indices, clipped index), formed on the host.
 * idx, w, table, out on the device; n_q <= 65535. */
int gm_ri_bilinear(gm_ctx* ctx, int n_mr, int n_mi, int64_t n_elem, const double* table,
                   int64_t n_q, const int32_t* idx, const double* w, double* out,
                   gm_stream stream);

/* P normalisation of six independent elements, pmat [e][6][n_ang] in the
 * order p11, p12, p22, p33, p34, p44 (kernel mode: p22 != p11, p44 != p33),
 * in place, and pback [e][6] in the table's order p11, p12, p33, p34, p22,
 * p44.  Replaces dointegration.py:978-988 for the kernel path. */
int gm_pmatrix_normalize6(gm_ctx* ctx, int64_t n_entries, int n_ang, const double* theta_rad,
                          double* pmat, double* pback, gm_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* GEOSMIE_H */
