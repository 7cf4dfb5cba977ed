// Correctly rounded acos and sincos for the RMHD quartic's resolvent cubic
// (Src/RMHD/rmhd.hpp:441-444: theta = acos(arg)/3, cos(theta), sin(theta)).
//
// The reference's CPU build takes these from the host libm (glibc), whose
// acos, sin and cos return the correctly rounded double for all but a tiny
// fraction of arguments.  The device math library is faithful (about 1 ulp),
// so its last bit differs often, and the RMHD state drifts from the
// reference's at the 1e-12 level.  These routines compute the value to about
// 2^-100 relative accuracy in double-double arithmetic and round it once, so
// they return the correctly rounded double unless the exact value lies within
// ~2^-100 of a rounding midpoint.  The same source runs on the host (tests
// compare it with glibc on 1e8 arguments) and on the device.
//
// Method (the classical table + polynomial evaluation, e.g. Muller,
// "Elementary Functions: Algorithms and Implementation", ch. 6 and 11, with
// error-free transformations from Dekker 1971 and Ogita, Rump & Oishi 2005):
//  - sincos(x), 0 <= x <= 2.19: x = t_k + r with t_k = k/64 and |r| <= 1/128
//    (r exact by Sterbenz); sin(t_k), cos(t_k) as double-doubles from the
//    table; sin r and cos r by their Taylor series to r^11 and r^10, the two
//    leading correction terms in double-double; the addition formulas in
//    double-double.
//  - acos(x): one Newton step in double-double from the library's acos of
//    |x|, whose error (a few ulp) becomes ~2^-104 after the step.
//    |x| <= 1/2: theta = theta0 + (cos theta0 - x)/sin theta0.  |x| > 1/2:
//    acos x = 2 phi (x > 0) or pi - 2 phi (x < 0) with sin^2 phi = h =
//    (1 - |x|)/2 (exact), phi = phi0 - (sin^2 phi0 - h)/sin(2 phi0),
//    phi0 = acos(|x|)/2, which keeps full relative accuracy as x -> +-1.
// Explicit fma() calls are the error-free products; they are exact on both
// sides regardless of -ffp-contract.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "crmath_table.hpp"

namespace gpl { namespace cr {

#define GPL_CR_HD __host__ __device__ __forceinline__

struct dd { double hi, lo; };

// Knuth's TwoSum: hi + lo == a + b exactly
GPL_CR_HD dd two_sum(double a, double b)
{
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
// Dekker's FastTwoSum, |a| >= |b|
GPL_CR_HD dd fast_two_sum(double a, double b)
{
  const double s = a + b;
  return {s, b - (s - a)};
}
// hi + lo == a*b exactly
GPL_CR_HD dd two_prod(double a, double b)
{
  const double p = a*b;
  return {p, fma(a, b, -p)};
}
GPL_CR_HD dd dd_add(dd a, dd b)
{
  dd s = two_sum(a.hi, b.hi);
  s.lo += a.lo + b.lo;
  return fast_two_sum(s.hi, s.lo);
}
GPL_CR_HD dd dd_mul(dd a, dd b)
{
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi*b.lo + a.lo*b.hi;
  return fast_two_sum(p.hi, p.lo);
}
GPL_CR_HD dd dd_mul_d(dd a, double b)
{
  dd p = two_prod(a.hi, b);
  p.lo += a.lo*b;
  return fast_two_sum(p.hi, p.lo);
}

GPL_CR_HD const double *table_row(int k)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return kSinCosDev[k];
#else
  return kSinCosHost[k];
#endif
}

// sin and cos of x, 0 <= x <= 2.19 (table index k <= 140), as double-doubles (~2^-100 relative for
// sin and cos away from their zeros; absolute ~2^-105 near cos's zero)
GPL_CR_HD void dd_sincos(double x, dd &s, dd &c)
{
  const int k = (int)(x*64.0 + 0.5);
  const double r = x - k*0.015625;               // exact
  const double zh = r*r, zl = fma(r, r, -zh);     // z = r^2 = zh + zl
  // sin r = r (1 + z (s1 + z (s2 + z (s3 + z (s4 + z s5)))))
  const double st = -1.0/5040.0 + zh*(1.0/362880.0 + zh*(-1.0/39916800.0));
  dd t = fast_two_sum(kS2Hi, zh*st);
  t.lo += kS2Lo;                                   // s2 + z st
  dd w = dd_mul(dd{zh, zl}, t);
  w = dd_add(dd{kS1Hi, kS1Lo}, w);                 // s1 + z (...)
  w = dd_mul(dd{zh, zl}, w);                       // z (...)
  const dd rw = dd_mul_d(w, r);                    // |r w| < 2^-16 |r|
  dd sr = fast_two_sum(r, rw.hi);
  sr.lo += rw.lo;                                  // sin r = r + r w
  // cos r = 1 + z (-1/2 + z (c2 + z (c3 + z (c4 + z c5)))), cw = cos r - 1
  const double ct = -1.0/720.0 + zh*(1.0/40320.0 + zh*(-1.0/3628800.0));
  dd u = fast_two_sum(kC2Hi, zh*ct);
  u.lo += kC2Lo;                                   // c2 + z ct
  u = dd_mul(dd{zh, zl}, u);
  u = dd_add(dd{-0.5, 0.0}, u);
  const dd cw = dd_mul(dd{zh, zl}, u);
  // addition formulas: sin x = S + S cw + C sin r, cos x = C + C cw - S sin r
  const double *row = table_row(k);
  const dd S{row[0], row[1]}, C{row[2], row[3]};
  s = dd_add(dd_add(S, dd_mul(S, cw)), dd_mul(C, sr));
  const dd msr{-sr.hi, -sr.lo};
  c = dd_add(dd_add(C, dd_mul(C, cw)), dd_mul(S, msr));
}

// sin(x) and cos(x) correctly rounded for 0 <= x <= 2.19 (the resolvent
// cubic's theta lies in [0, pi/3]); other arguments (NaN) take the library's.
GPL_CR_HD void sincos_cr(double x, double *sn, double *cs)
{
  if (!(x >= 0.0 && x <= 2.19)) { sincos(x, sn, cs); return; }
  dd s, c;
  dd_sincos(x, s, c);
  *sn = s.hi + s.lo;
  *cs = c.hi + c.lo;
}

// The Newton step of acos_cr for -1 < x < 1 from its seed (acos_seed).  A
// seed a few ulp off still gives the correctly rounded result (the step
// squares its relative error).  Both forms of the step are evaluated from one
// double-double sincos of the seed and the right one is selected, so lanes of
// a wavefront that take different forms do not run the expensive part twice.
//  |x| <= 1/2: theta = theta0 + (cos theta0 - x)/sin theta0, theta0 in [pi/3, 2pi/3];
//  |x| >  1/2: acos x = 2 phi (x > 0) or pi - 2 phi (x < 0) with sin^2 phi = h =
//              (1 - |x|)/2 (exact), phi = phi0 - (sin^2 phi0 - h)/sin(2 phi0),
//              phi0 in (0, pi/6], which keeps full relative accuracy as x -> +-1.
GPL_CR_HD double acos_refine(double x, double seed)
{
  dd s, c;
  dd_sincos(seed, s, c);
  const double d1 = (c.hi - x) + c.lo;             // cos t0 - x
  const double r_mid = seed + d1/s.hi;
  const double h = (1.0 - fabs(x))*0.5;            // sin^2 phi, exact
  const dd s2 = dd_mul(s, s);
  const double d2 = (s2.hi - h) + s2.lo;           // sin^2 phi0 - h
  const double delta = -d2/(2.0*s.hi*c.hi);
  const double r_pos = 2.0*(seed + delta);         // exact doubling of RN(phi)
  dd t = two_sum(kPiHi, -2.0*seed);                // pi - 2 phi
  t.lo += kPiLo - 2.0*delta;
  const double r_neg = t.hi + t.lo;
  return fabs(x) <= 0.5 ? r_mid : (x > 0.0 ? r_pos : r_neg);
}

// The seed of the Newton step from one library acos of |x| (a few ulp):
// theta0 = acos x for |x| <= 1/2, phi0 = acos(|x|)/2 otherwise.
GPL_CR_HD double acos_seed(double x)
{
  const double t = acos(fabs(x));
  if (fabs(x) <= 0.5) return x >= 0.0 ? t : kPiHi - t;
  return 0.5*t;
}

// acos(x) correctly rounded for -1 <= x <= 1; NaN takes the library's.
GPL_CR_HD double acos_cr(double x)
{
  if (!(x >= -1.0 && x <= 1.0)) return acos(x);
  if (x == 1.0) return 0.0;
  if (x == -1.0) return kPiHi;
  return acos_refine(x, acos_seed(x));
}

#undef GPL_CR_HD

}}  // namespace gpl::cr
