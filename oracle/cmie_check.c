/*
 * Driver for the sanitizer build of the C restatement (TEST INFRASTRUCTURE
 * ONLY): runs cmie_batch (homogeneous, coated literal and stable forms, the
 * dispatch cases of mie_coeffs.py:61-71) and cmie_chain_homog on small
 * batches that reach every code path -- Rayleigh-limit and large sizes,
 * dense and absorbing shells, points needing the start correction -- on
 * several OpenMP threads, and checks that every output is finite.  Built
 * with -fsanitize=address,undefined by `make -C oracle sanitize`; a memory
 * or undefined-behaviour error aborts with a report.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int cmie_batch(int64_t n, const double* x, const double* y, const double* m, const double* m2,
               double* props, int threads, int stable);
int cmie_chain_homog(int64_t n, const double* x, const double* m, const double* r,
                     const int32_t* row, const double* w, int n_rows, int n_ang,
                     const double* pt, int nmax_cap, double* pm, double* sc, int threads);

static int all_finite(const double* v, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(v[i])) return 0;
  return 1;
}

int main(void) {
  enum { N = 96 };
  double x[N], y[N], m[2 * N], m2[2 * N], props[6 * N];
  for (int i = 0; i < N; ++i) {
    double s = pow(10.0, -2.0 + 4.3 * i / (N - 1));      /* 0.01 .. ~200 */
    y[i] = s;
    x[i] = (i % 7 == 0) ? 0.0 : (i % 11 == 0 ? s : s * (0.3 + 0.65 * (i % 5) / 4.0));
    m[2 * i] = (i % 2) ? 1.95 : 1.33;
    m[2 * i + 1] = (i % 2) ? 0.79 : 1e-4;
    m2[2 * i] = (i % 3 == 0) ? 1.5 : 1.45;
    m2[2 * i + 1] = (i % 3 == 1) ? 0.5 : 1e-3;
  }
  int ok = 1;
  for (int stable = 0; stable < 2; ++stable) {
    cmie_batch(N, x, y, m, m2, props, 3, stable);
    ok &= all_finite(props, 6 * N);
  }
  cmie_batch(N, y, NULL, m, NULL, props, 2, 0);
  ok &= all_finite(props, 6 * N);
  /* the homogeneous chain: two rows, 13 angles */
  enum { NA = 13 };
  const int cap = 300;
  double* pt = (double*)calloc((size_t)2 * cap * NA, sizeof(double));
  for (int a = 0; a < NA; ++a) {
    double u = cos(M_PI * a / (NA - 1));
    double p0 = 0.0, p1 = 1.0;
    for (int n = 1; n <= cap; ++n) {
      double p = (n == 1) ? 1.0 : ((2.0 * n - 1.0) / (n - 1.0)) * u * p1 - (n / (n - 1.0)) * p0;
      if (n > 1) { p0 = p1; p1 = p; }
      double t = n * u * p1 - (n + 1.0) * p0;
      double f = (2.0 * n + 1.0) / (n * (n + 1.0));
      pt[(size_t)(n - 1) * NA + a] = f * p1;
      pt[(size_t)cap * NA + (size_t)(n - 1) * NA + a] = f * t;
    }
  }
  double r[N], w[N], pm[2 * 4 * NA], sc[2 * 4];
  int32_t row[N];
  for (int i = 0; i < N; ++i) {
    r[i] = y[i] * 0.55 / (2 * M_PI);
    w[i] = 1.0 / (1.0 + i);
    row[i] = i % 2;
  }
  int rc = cmie_chain_homog(N, y, m, r, row, w, 2, NA, pt, cap, pm, sc, 4);
  ok &= (rc == 0) && all_finite(pm, 2 * 4 * NA) && all_finite(sc, 8);
  free(pt);
  printf(ok ? "cmie sanitizer check ok\n" : "cmie sanitizer check: non-finite output\n");
  return ok ? 0 : 1;
}
