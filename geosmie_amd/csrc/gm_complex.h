// FP64 complex arithmetic for the gfx950 Mie kernels.
//
// Plain structs (no std::complex) so the compiler keeps every value in VGPRs
// and fuses multiply-adds.  Division is branch-free: the divisor is scaled by
// a power of two (exponent only, exact) so |b|^2 cannot overflow or
// underflow, the reciprocal comes from v_rcp_f64 refined by two Newton steps,
// and the scale is undone; quotients are within a few ulps of numpy's
// complex128 division.  Recurrences that need more use the double-double
// reciprocals further down.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

struct cd {
  double re, im;
};

__device__ __forceinline__ cd mk(double re, double im) { return cd{re, im}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator-(cd a) { return cd{-a.re, -a.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
  return cd{fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ cd operator*(cd a, double s) { return cd{a.re * s, a.im * s}; }
__device__ __forceinline__ cd operator*(double s, cd a) { return cd{a.re * s, a.im * s}; }
__device__ __forceinline__ cd operator+(cd a, double s) { return cd{a.re + s, a.im}; }
__device__ __forceinline__ cd operator-(cd a, double s) { return cd{a.re - s, a.im}; }
__device__ __forceinline__ cd operator-(double s, cd a) { return cd{s - a.re, -a.im}; }
__device__ __forceinline__ cd conj(cd a) { return cd{a.re, -a.im}; }
__device__ __forceinline__ double norm2(cd a) { return fma(a.re, a.re, a.im * a.im); }
__device__ __forceinline__ double cabs_(cd a) { return hypot(a.re, a.im); }

// 1/d to ~1 ulp for normal-range d: the v_rcp_f64 estimate refined by two
// Newton steps (5 instructions; the IEEE division sequence is ~10 with
// data-dependent fixups).
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

// a / b without branches: b is scaled by an exact power of two so that
// max(|re|, |im|) lies in [0.5, 1), then a conj(b') / |b'|^2 and the scale is
// undone.  Overflow-safe like Smith's method (numpy's), a few ulps from it,
// and free of the divergent branch Smith's method takes on |re| >= |im|.
__device__ __forceinline__ cd cdiv(cd a, cd b) {
  const double mx = fmax(fabs(b.re), fabs(b.im));
  const int e = __builtin_amdgcn_frexp_exp(mx);
  const double br = ldexp(b.re, -e), bi = ldexp(b.im, -e);
  const double inv = rcp_nr(fma(br, br, bi * bi));
  const double qr = fma(a.re, br, a.im * bi) * inv;
  const double qi = fma(a.im, br, -(a.re * bi)) * inv;
  return cd{ldexp(qr, -e), ldexp(qi, -e)};
}
__device__ __forceinline__ cd cdiv(double a, cd b) { return cdiv(cd{a, 0.0}, b); }

// a / b and c / d sharing one reciprocal: each divisor scaled by its own
// power of two as in cdiv, then 1 / (|b'|^2 |d'|^2) (the product lies in
// [1/16, 4)) split into 1/|b'|^2 = |d'|^2 r and 1/|d'|^2 = |b'|^2 r.
__device__ __forceinline__ void cdiv2(cd a, cd b, cd c, cd d, cd& q1, cd& q2) {
  const int e1 = __builtin_amdgcn_frexp_exp(fmax(fabs(b.re), fabs(b.im)));
  const int e2 = __builtin_amdgcn_frexp_exp(fmax(fabs(d.re), fabs(d.im)));
  const double br = ldexp(b.re, -e1), bi = ldexp(b.im, -e1);
  const double dr = ldexp(d.re, -e2), di = ldexp(d.im, -e2);
  const double nb = fma(br, br, bi * bi), nd = fma(dr, dr, di * di);
  const double r = rcp_nr(nb * nd);
  const double ib = nd * r, id = nb * r;
  q1 = cd{ldexp(fma(a.re, br, a.im * bi) * ib, -e1), ldexp(fma(a.im, br, -(a.re * bi)) * ib, -e1)};
  q2 = cd{ldexp(fma(c.re, dr, c.im * di) * id, -e2), ldexp(fma(c.im, dr, -(c.re * di)) * id, -e2)};
}

// 1 / t for the log-derivative recurrence.  |t| stays far from the FP64
// overflow/underflow range there (t = D_n + n/z), so no scaling is needed.
__device__ __forceinline__ cd crcp_fast(cd t) {
  double inv = rcp_nr(norm2(t));
  return cd{t.re * inv, -t.im * inv};
}

// Reciprocals carried as an unevaluated sum hi + lo (double-double).
//
// Recurrences multiply by 1/z, 1/m, 1/x at every step.  A reciprocal rounded
// once is a fixed relative error of up to an ulp, the same at every step --
// equivalent to perturbing z (or m) by an ulp, which near a Mie resonance
// moves Q by ~1e-9.  The reference divides at every step instead, whose
// roundings are independent.  a * (hi + lo) rounded once restores that.
struct ddr {
  double hi, lo;
};
struct cdd {
  cd hi, lo;
};
__device__ __forceinline__ ddr rcp_dd(double x) {
  double hi = 1.0 / x;
  double e = fma(-x, hi, 1.0);  // exact residual 1 - x hi
  return ddr{hi, e * hi};
}
__device__ __forceinline__ double mul_dd(double a, ddr b) { return fma(a, b.hi, a * b.lo); }
__device__ __forceinline__ cd mul_dd(cd a, ddr b) {
  return cd{fma(a.re, b.hi, a.re * b.lo), fma(a.im, b.hi, a.im * b.lo)};
}
// a / z for complex z: hi = 1/z (cdiv), lo = hi * (1 - z hi) with the
// residual formed from FMA-exact products.
__device__ __forceinline__ cdd crcp_dd(cd z) {
  cd hi = cdiv(cd{1.0, 0.0}, z);
  // residual r = 1 - z*hi, each product's rounding error recovered by fma
  double p1 = z.re * hi.re, e1 = fma(z.re, hi.re, -p1);
  double p2 = z.im * hi.im, e2 = fma(z.im, hi.im, -p2);
  double q1 = z.re * hi.im, f1 = fma(z.re, hi.im, -q1);
  double q2 = z.im * hi.re, f2 = fma(z.im, hi.re, -q2);
  double rr = ((1.0 - p1) + p2) - e1 + e2;
  double ri = -(q1 + q2) - f1 - f2;
  cd r = cd{rr, ri};
  return cdd{hi, hi * r};
}
// a * (hi + lo) with a real scalar a (e.g. (j+1)/z)
__device__ __forceinline__ cd mul_dd(double a, cdd b) {
  return cd{fma(a, b.hi.re, a * b.lo.re), fma(a, b.hi.im, a * b.lo.im)};
}
// a * (hi + lo) with complex a
__device__ __forceinline__ cd mul_dd(cd a, cdd b) { return a * b.hi + a * b.lo; }

// Principal square root, the branch numpy's complex sqrt returns.
__device__ __forceinline__ cd csqrt_(cd z) {
  double a = z.re, b = z.im;
  if (a == 0.0 && b == 0.0) return cd{0.0, b};
  double r = hypot(a, b);
  if (a >= 0.0) {
    double t = sqrt(0.5 * (r + a));
    return cd{t, b / (2.0 * t)};
  }
  double t = sqrt(0.5 * (r - a));
  return cd{fabs(b) / (2.0 * t), copysign(t, b)};
}

__device__ __forceinline__ cd csin_(cd z) {
  double s, c;
  sincos(z.re, &s, &c);
  return cd{s * cosh(z.im), c * sinh(z.im)};
}
__device__ __forceinline__ cd ccos_(cd z) {
  double s, c;
  sincos(z.re, &s, &c);
  return cd{c * cosh(z.im), -s * sinh(z.im)};
}

// sin z and cos z together: one sincos of Re z, and sinh, cosh of Im z from
// one expm1: with m = e^|t| - 1, sinh |t| = (m + m/(m+1))/2 and
// cosh t = (m + 1 + 1/(m+1))/2, free of cancellation for small |t|.
__device__ __forceinline__ void csincos_(cd z, cd& sn, cd& cs) {
  double s, c;
  sincos(z.re, &s, &c);
  const double m = expm1(fabs(z.im));
  const double ie = 1.0 / (m + 1.0);
  const double mi = m < 1e300 ? m * ie : 1.0;   // m / (m + 1), also where e^|t| overflows
  const double sh = copysign(0.5 * (m + mi), z.im);
  const double ch = 0.5 * ((m + 1.0) + ie);
  sn = cd{s * ch, c * sh};
  cs = cd{c * ch, -s * sh};
}
