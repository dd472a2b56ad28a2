/*
 * C restatement of the pymiecoated Mie algorithm (homogeneous and coated
 * spheres, coefficients and efficiencies), FP64, OpenMP over points.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a fast oracle at sizes where
 * the numpy/scipy restatement (oracle/mie.py) is too slow, and by bench.py
 * as the timed host-CPU baseline.  Never linked into the product.
 *
 * Follows src/pymiecoated/pymiecoated/mie_coeffs.py:
 *   n_max, nmx rules          :99, :101, :203-205 (half-to-even rounding)
 *   D_n downward recurrence   :118-121, :214-218 (start D_nmx = 0)
 *   homogeneous a_n, b_n      :123-128
 *   coated shell combination  :230-249
 * and src/pymiecoated/pymiecoated/mie_props.py:28-70 for the efficiencies.
 *
 * The reference evaluates psi_n, chi_n with scipy jv/yv (mie_coeffs.py:156-161,
 * 223-224).  This restatement computes them by recurrence instead: chi_n
 * upward; psi_n upward for n <= |z| and, above, psi_n = psi_{n-1}/(D_n(z)+n/z)
 * with D_n(z) from a downward recurrence started 16 + 8 ceil(|z|^(1/3)) above
 * n_max.  Pinned against the scipy-based oracle/mie.py and the mpmath oracle
 * in tests/test_oracle_cmie.py.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static int nmax_rule(double s) { return (int)rint(2.0 + s + 4.0 * pow(s, 1.0 / 3.0)); }
static int nmx_rule(int nmax, double zabs) { return (int)rint(fmax((double)nmax, zabs) + 16.0); }
static int ratio_start(int nmax, double zabs) { return nmax + 16 + 8 * (int)ceil(cbrt(zabs)); }

/* Smith's complex division (numpy's algorithm) */
static cplx cdivs(cplx a, cplx b) {
  double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
  if (fabs(br) >= fabs(bi)) {
    if (br == 0.0 && bi == 0.0) return CMPLX(ar / 0.0, ai / 0.0);
    double r = bi / br, s = 1.0 / (br + bi * r);
    return CMPLX((ar + ai * r) * s, (ai - ar * r) * s);
  }
  double r = br / bi, s = 1.0 / (bi + br * r);
  return CMPLX((ar * r + ai) * s, (ai * r - ar) * s);
}

static cplx csqrt_p(cplx z) { return csqrt(z); }

/* D_1..D_nmax of z from D_nmx = 0 into d[0..nmax-1] */
static void dn_down(cplx z, int nmx, int nmax, cplx* d) {
  cplx dd = 0.0;
  for (int j = nmx - 1; j >= 1; --j) {
    cplx r = (double)(j + 1) / z;
    dd = r - cdivs(1.0, dd + r);
    if (j <= nmax) d[j - 1] = dd;
  }
}

/* psi_1..psi_nmax (index 0..nmax-1) and chi_0..chi_nmax of complex z */
static void riccati(cplx z, int nmax, cplx* psi, cplx* chi, cplx* tmp) {
  int nsw = (int)cabs(z);
  if (nsw > nmax) nsw = nmax;
  if (nsw < nmax) {
    cplx dd = 0.0;
    for (int n = ratio_start(nmax, cabs(z)); n > nsw; --n) {
      cplx r = (double)n / z;
      cplx t = dd + r;
      if (n <= nmax) tmp[n - 1] = t;
      dd = r - cdivs(1.0, t);
    }
  }
  cplx pp = ccos(z), p = csin(z), cpp = -csin(z), c = ccos(z);
  chi[0] = c;
  for (int n = 1; n <= nmax; ++n) {
    cplx f = (double)(2 * n - 1) / z;
    cplx pn = (n <= nsw) ? f * p - pp : cdivs(p, tmp[n - 1]);
    cplx cn = f * c - cpp;
    psi[n - 1] = pn;
    chi[n] = cn;
    pp = p; p = pn; cpp = c; c = cn;
  }
}

static void props_of(const cplx* an, const cplx* bn, int nmax, double y, double* out) {
  double sext = 0, ssca = 0, asy1 = 0, asy2 = 0;
  cplx qb = 0.0;
  for (int k = 1; k <= nmax; ++k) {
    cplx a = an[k - 1], b = bn[k - 1];
    double dn = k, cn = 2 * dn + 1;
    sext += cn * (creal(a) + creal(b));
    ssca += cn * (creal(a) * creal(a) + cimag(a) * cimag(a) + creal(b) * creal(b) + cimag(b) * cimag(b));
    qb += (a - b) * ((k & 1) ? -cn : cn);
    if (k < nmax) {
      cplx a1 = an[k], b1 = bn[k];
      asy1 += dn * (dn + 2) / (dn + 1) *
              (creal(a) * creal(a1) + cimag(a) * cimag(a1) + creal(b) * creal(b1) + cimag(b) * cimag(b1));
    }
    asy2 += cn / (dn * (dn + 1)) * (creal(a) * creal(b) + cimag(a) * cimag(b));
  }
  double y2 = y * y;
  double qext = 2 * sext / y2, qsca = 2 * ssca / y2, qbv = (creal(qb) * creal(qb) + cimag(qb) * cimag(qb)) / y2;
  out[0] = qext; out[1] = qsca; out[2] = qext - qsca; out[3] = qbv;
  out[4] = 4 / y2 * (asy1 + asy2) / qsca; out[5] = qbv / qsca;
}

/* Homogeneous sphere (mie_coeffs.py:83-130) with mu = 1. */
static int homog(double x, cplx m, cplx* an, cplx* bn, cplx* work) {
  cplx z = m * x;
  int nmax = nmax_rule(x);
  int nmx = nmx_rule(nmax, cabs(z));
  cplx* D = work;
  cplx* psi = work + nmax;
  cplx* chi = psi + nmax;
  cplx* tmp = chi + nmax + 1;
  dn_down(z, nmx, nmax, D);
  riccati(x, nmax, psi, chi, tmp);
  double sx = sin(x);
  for (int n = 1; n <= nmax; ++n) {
    double p = creal(psi[n - 1]), p1 = (n == 1) ? sx : creal(psi[n - 2]);
    double c = creal(chi[n]), c1 = creal(chi[n - 1]);
    cplx xi = CMPLX(p, -c), xi1 = CMPLX(p1, -c1);
    cplx da = cdivs(D[n - 1], m) + n / x;
    cplx db = D[n - 1] * m + n / x;
    an[n - 1] = cdivs(da * p - p1, da * xi - xi1);
    bn[n - 1] = cdivs(db * p - p1, db * xi - xi1);
  }
  return nmax;
}

/* Coated sphere (mie_coeffs.py:183-251). */
static int coated(double x, double y, cplx m1, cplx m2, cplx* an, cplx* bn, cplx* work) {
  cplx m = m2 / m1, u = m1 * x, v = m2 * x, w = m2 * y;
  int nmax = nmax_rule(y);
  double mx = fmax(cabs(m1 * y), cabs(w));
  int nmx = nmx_rule(nmax, mx);
  cplx* dnu = work;
  cplx* dnv = dnu + nmax;
  cplx* dnw = dnv + nmax;
  cplx* pv = dnw + nmax;
  cplx* chv = pv + nmax;
  cplx* pw = chv + nmax + 1;
  cplx* chw = pw + nmax;
  cplx* py = chw + nmax + 1;
  cplx* chy = py + nmax;
  cplx* tmp = chy + nmax + 1;
  dn_down(u, nmx, nmax, dnu);
  dn_down(v, nmx, nmax, dnv);
  dn_down(w, nmx, nmax, dnw);
  riccati(v, nmax, pv, chv, tmp);
  riccati(w, nmax, pw, chw, tmp);
  riccati(y, nmax, py, chy, tmp);
  double sy = sin(y);
  for (int n = 1; n <= nmax; ++n) {
    int i = n - 1;
    cplx uu = m * dnu[i] - dnv[i];
    cplx vv = cdivs(dnu[i], m) - dnv[i];
    cplx PV = pv[i], CHV = chv[n], PW = pw[i], CHW = chw[n];
    cplx fv = cdivs(PV, CHV);
    cplx ku1 = cdivs(uu * fv, PW), kv1 = cdivs(vv * fv, PW);
    cplx pt = PW - CHW * fv;
    cplx prat = cdivs(cdivs(PW, PV), CHV);
    cplx ku2 = uu * pt + prat, kv2 = vv * pt + prat;
    cplx dns = cdivs(ku1, ku2) + dnw[i];
    cplx gns = cdivs(kv1, kv2) + dnw[i];
    double nrat = n / y;
    cplx a1 = cdivs(dns, m2) + nrat, b1 = m2 * gns + nrat;
    double p = creal(py[i]), p1 = (n == 1) ? sy : creal(py[i - 1]);
    double c = creal(chy[n]), c1 = creal(chy[n - 1]);
    cplx gsy = CMPLX(p, -c), gs1y = CMPLX(p1, -c1);
    an[i] = cdivs(p * a1 - p1, gsy * a1 - gs1y);
    bn[i] = cdivs(p * b1 - p1, gsy * b1 - gs1y);
  }
  return nmax;
}

/*
 * Coated sphere, the same coefficients rearranged to avoid the reference's
 * divisions by psi_n(m2 y) (ku1 = uu fv / pw, mie_coeffs.py:234) and its
 * products D_n(m2 x) psi_n(m2 x) near zeros of psi_n(m2 x), where any FP64
 * evaluation of the literal formula loses up to ~|log10 psi| digits.
 * With psi' = psi_{n-1} - (n/z) psi_n, chi' likewise, and the Wronskian
 * psi' chi - psi chi' = 1:
 *   uu psi_v = m D_u psi_v - psi'_v - dD_v psi_v,   (dD = D_ref - D_exact)
 *   dns = [psi'_w + uu psi_v (psi'_w chi_v - chi'_w psi_v)]
 *       / [psi_w + uu psi_v (psi_w chi_v - chi_w psi_v)] + dD_w
 * which equals dns1 + dnw of mie_coeffs.py:230-244 exactly.  dD_v, dD_w are
 * zero unless the reference's start index is too close to |v| or |w| for
 * D_n to have converged (then both recurrences are run and differenced).
 */
static int converged_start(int nmx, double zabs) { return (double)nmx - zabs >= 8.0 * cbrt(zabs); }

static int coated_stable(double x, double y, cplx m1, cplx m2, cplx* an, cplx* bn, cplx* work) {
  cplx m = m2 / m1, u = m1 * x, v = m2 * x, w = m2 * y;
  int nmax = nmax_rule(y);
  double mx = fmax(cabs(m1 * y), cabs(w));
  int nmx = nmx_rule(nmax, mx);
  cplx* dnu = work;
  cplx* ddv = dnu + nmax;
  cplx* ddw = ddv + nmax;
  cplx* pv = ddw + nmax;
  cplx* chv = pv + nmax;
  cplx* pw = chv + nmax + 1;
  cplx* chw = pw + nmax;
  cplx* py = chw + nmax + 1;
  cplx* chy = py + nmax;
  cplx* t1 = chy + nmax + 1;
  cplx* t2 = t1 + nmax;
  cplx* tmp = t2 + nmax;
  dn_down(u, nmx, nmax, dnu);
  cplx zs[2] = {v, w};
  cplx* dd[2] = {ddv, ddw};
  for (int q = 0; q < 2; ++q) {
    double za = cabs(zs[q]);
    if (converged_start(nmx, za)) {
      for (int i = 0; i < nmax; ++i) dd[q][i] = 0.0;
    } else {
      dn_down(zs[q], nmx, nmax, t1);
      dn_down(zs[q], nmx + 16 + 8 * (int)ceil(cbrt(za)), nmax, t2);
      for (int i = 0; i < nmax; ++i) dd[q][i] = t1[i] - t2[i];
    }
  }
  riccati(v, nmax, pv, chv, tmp);
  riccati(w, nmax, pw, chw, tmp);
  riccati(y, nmax, py, chy, tmp);
  cplx pv0 = csin(v), pw0 = csin(w);
  double sy = sin(y);
  for (int n = 1; n <= nmax; ++n) {
    int i = n - 1;
    cplx PV = pv[i], CHV = chv[n], PW = pw[i], CHW = chw[n];
    cplx pv1 = (n == 1) ? pv0 : pv[i - 1];
    cplx pw1 = (n == 1) ? pw0 : pw[i - 1];
    cplx dpv = pv1 - (double)n / v * PV;
    cplx dpw = pw1 - (double)n / w * PW;
    cplx dcw = chw[n - 1] - (double)n / w * CHW;
    cplx U = m * dnu[i] * PV - dpv - ddv[i] * PV;
    cplx V = cdivs(dnu[i], m) * PV - dpv - ddv[i] * PV;
    cplx A = PW * CHV - CHW * PV, B = dpw * CHV - dcw * PV;
    cplx dns = cdivs(dpw + U * B, PW + U * A) + ddw[i];
    cplx gns = cdivs(dpw + V * B, PW + V * A) + ddw[i];
    double nrat = n / y;
    cplx a1 = cdivs(dns, m2) + nrat, b1 = m2 * gns + nrat;
    double p = creal(py[i]), p1 = (n == 1) ? sy : creal(py[i - 1]);
    double c = creal(chy[n]), c1 = creal(chy[n - 1]);
    cplx gsy = CMPLX(p, -c), gs1y = CMPLX(p1, -c1);
    an[i] = cdivs(p * a1 - p1, gsy * a1 - gs1y);
    bn[i] = cdivs(p * b1 - p1, gsy * b1 - gs1y);
  }
  return nmax;
}

/*
 * Batch entry: props[6*n] planes (qext, qsca, qabs, qb, asy, qratio).
 * x[n], m[2n] interleaved; y == NULL -> homogeneous, else m2[2n] shell.
 * threads <= 0: OpenMP default.  Returns 0.
 */
int cmie_batch(int64_t n, const double* x, const double* y, const double* m, const double* m2,
               double* props, int threads, int stable) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
  {
    int cap = 0;
    cplx* work = NULL;
    cplx* an = NULL;
    cplx* bn = NULL;
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
      double s = y ? y[i] : x[i];
      int nm = nmax_rule(s);
      int need = 16 * nm + 64 + ratio_start(nm, 3.0 * s + 10.0);
      if (need > cap) {
        free(work); free(an); free(bn);
        cap = need * 2;
        work = (cplx*)malloc(sizeof(cplx) * cap);
        an = (cplx*)malloc(sizeof(cplx) * cap);
        bn = (cplx*)malloc(sizeof(cplx) * cap);
      }
      cplx mi = CMPLX(m[2 * i], m[2 * i + 1]);
      double out[6];
      if (!y) {
        int k = homog(x[i], mi, an, bn, work);
        props_of(an, bn, k, x[i], out);
      } else {
        cplx mo = CMPLX(m2[2 * i], m2[2 * i + 1]);
        int k;
        if (x[i] == y[i] || mi == mo) k = homog(y[i], mi, an, bn, work);
        else if (x[i] == 0.0) k = homog(y[i], mo, an, bn, work);
        else if (stable && fabs(cimag(mo)) * y[i] <= 1.0)
          k = coated_stable(x[i], y[i], mi, mo, an, bn, work);
        else k = coated(x[i], y[i], mi, mo, an, bn, work);
        props_of(an, bn, k, y[i], out);
      }
      for (int q = 0; q < 6; ++q) props[q * n + i] = out[q];
    }
    free(work); free(an); free(bn);
  }
  return 0;
}

/*
 * Homogeneous runoptics chain on the host (the CPU baseline of the plain
 * homogeneous workload): per point the coefficients (homog above), S1/S2 at
 * every angle from shared pi'/tau' tables (mie_props.py:133-149 with the
 * tables of mie_coated.py:160-179), the Mueller elements s11, s12, s33, s34
 * (dointegration.py:1045-1050) and their number-weighted sums per row
 * (dointegration.py:1164-1166), plus the area-weighted efficiency sums
 * sum w pi r^2, sum w pi r^2 Qext, sum w pi r^2 Qsca, sum w pi r^2 Qsca g
 * (dointegration.py:1178-1190).  Each point belongs to one row (row[i]) with
 * weight w[i] and radius r[i].  OpenMP over points with per-thread row
 * accumulators added in thread order at the end.
 *   pt[2 * nmax_cap * n_ang]: pi' then tau', [n][angle]
 *   pm[n_rows * 4 * n_ang], sc[n_rows * 4] (zeroed here)
 */
int cmie_chain_homog(int64_t n, const double* x, const double* m, const double* r,
                     const int32_t* row, const double* w, int n_rows, int n_ang,
                     const double* pt, int nmax_cap, double* pm, double* sc, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  const double* pi_t = pt;
  const double* tau_t = pt + (size_t)nmax_cap * n_ang;
  for (int64_t k = 0; k < (int64_t)n_rows * 4 * n_ang; ++k) pm[k] = 0.0;
  for (int k = 0; k < n_rows * 4; ++k) sc[k] = 0.0;
  int bad = 0;
#pragma omp parallel
  {
    double* acc = (double*)calloc((size_t)n_rows * 4 * n_ang, sizeof(double));
    double* asc = (double*)calloc((size_t)n_rows * 4, sizeof(double));
    cplx* work = (cplx*)malloc(sizeof(cplx) * (size_t)(16 * nmax_cap + 64 + ratio_start(nmax_cap, 4.0 * nmax_cap)));
    cplx* an = (cplx*)malloc(sizeof(cplx) * (size_t)nmax_cap);
    cplx* bn = (cplx*)malloc(sizeof(cplx) * (size_t)nmax_cap);
    cplx* s1 = (cplx*)malloc(sizeof(cplx) * (size_t)n_ang);
    cplx* s2 = (cplx*)malloc(sizeof(cplx) * (size_t)n_ang);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
      if (nmax_rule(x[i]) > nmax_cap) { bad = 1; continue; }
      cplx mi = CMPLX(m[2 * i], m[2 * i + 1]);
      int k = homog(x[i], mi, an, bn, work);
      double out[6];
      props_of(an, bn, k, x[i], out);
      for (int j = 0; j < n_ang; ++j) { s1[j] = 0.0; s2[j] = 0.0; }
      for (int q = 0; q < k; ++q) {
        const double* p = pi_t + (size_t)q * n_ang;
        const double* t = tau_t + (size_t)q * n_ang;
        cplx a = an[q], b = bn[q];
        for (int j = 0; j < n_ang; ++j) {
          s1[j] += a * p[j] + b * t[j];
          s2[j] += a * t[j] + b * p[j];
        }
      }
      double wi = w[i];
      double* o = acc + (size_t)row[i] * 4 * n_ang;
      for (int j = 0; j < n_ang; ++j) {
        double a1 = creal(s1[j]) * creal(s1[j]) + cimag(s1[j]) * cimag(s1[j]);
        double a2 = creal(s2[j]) * creal(s2[j]) + cimag(s2[j]) * cimag(s2[j]);
        cplx c = s1[j] * conj(s2[j]);
        o[j] += wi * 0.5 * (a1 + a2);
        o[n_ang + j] += wi * 0.5 * (a2 - a1);
        o[2 * n_ang + j] += wi * creal(c);
        o[3 * n_ang + j] += wi * cimag(c);
      }
      double area = wi * M_PI * r[i] * r[i];
      double* so = asc + (size_t)row[i] * 4;
      so[0] += area;
      so[1] += area * out[0];
      so[2] += area * out[1];
      so[3] += area * out[1] * out[4];
    }
#pragma omp for ordered schedule(static, 1)
    for (int tid = 0; tid < omp_get_num_threads(); ++tid) {
#pragma omp ordered
      {
        for (int64_t q = 0; q < (int64_t)n_rows * 4 * n_ang; ++q) pm[q] += acc[q];
        for (int q = 0; q < n_rows * 4; ++q) sc[q] += asc[q];
      }
    }
    free(acc); free(asc); free(work); free(an); free(bn); free(s1); free(s2);
  }
  return bad ? -1 : 0;
}
