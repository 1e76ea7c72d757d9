// Deterministic double-precision log / sin / cos shared by the host and gfx950.
//
// The transport histories consume natural logs (collision distance
// src/particle.cpp:243, log-grid bucket src/material.cpp:834, free gas
// src/physics.cpp:1008-1016, Maxwell/Watt src/random_dist.cpp:31) and sin/cos
// (rotate_angle src/math_functions.cpp:775-776, isotropic_direction
// src/distribution_multi.cpp:110-111, src/random_dist.cpp:28).  The reference
// takes them from glibc's libm.  A GPU cannot call glibc, and the histories must
// agree bit for bit between the host restatement and the kernels, so both sides
// use this header.  Every function is built only from IEEE-754 +, -, *, /, fma
// and integer bit operations, evaluated in double-double arithmetic to about
// 2^-80 relative error, and then rounded once.  The result is therefore the
// correctly rounded value except when the exact value lies within ~2^-80 of a
// rounding boundary; the tests measure how often that, or a glibc misrounding,
// makes this library differ from glibc.
//
// Methods (published, restated here):
//  * log: table-driven reduction x = 2^e * z, z*r_i = 1 + t with an exact
//    double-double t (Tang, "Table-driven implementation of the logarithm
//    function in IEEE floating-point arithmetic", ACM TOMS 16(4), 1990),
//    log1p(t) by its Taylor series with the first terms in double-double
//    (Dekker 1971 / Bailey QD double-double arithmetic).
//  * sin/cos: Cody-Waite reduction by pi/2 split in three 33-bit parts
//    (Cody & Waite, "Software Manual for the Elementary Functions", 1980),
//    then y = j/64 + z with tabulated sin(j/64), cos(j/64) in double-double
//    and Taylor series for sin(z), cos(z).
//
// Compile with -ffp-contract=off on both the host and the device so that no
// expression is contracted differently on the two sides.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DETM_HD __host__ __device__ inline
#else
#define DETM_HD inline
#endif

#include "openmc_amd/detmath_tables.h"

namespace detm {

struct dd {
  double hi;
  double lo;
};

DETM_HD uint64_t as_bits(double x)
{
  uint64_t u;
  __builtin_memcpy(&u, &x, 8);
  return u;
}

DETM_HD double from_bits(uint64_t u)
{
  double x;
  __builtin_memcpy(&x, &u, 8);
  return x;
}

DETM_HD double fma_(double a, double b, double c)
{
  return __builtin_fma(a, b, c);
}

// Knuth two-sum: s + e == a + b exactly.
DETM_HD dd two_sum(double a, double b)
{
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

// Dekker fast two-sum, valid for |a| >= |b| (or a == 0).
DETM_HD dd fast_two_sum(double a, double b)
{
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}

DETM_HD dd two_prod(double a, double b)
{
  double p = a * b;
  double e = fma_(a, b, -p);
  return {p, e};
}

DETM_HD dd dd_add(dd a, dd b)
{
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}

DETM_HD dd dd_mul(dd a, dd b)
{
  dd p = two_prod(a.hi, b.hi);
  double cross = a.hi * b.lo;
  double cross2 = a.lo * b.hi;
  p.lo += cross + cross2;
  return fast_two_sum(p.hi, p.lo);
}

DETM_HD dd dd_mul_d(dd a, double b)
{
  dd p = two_prod(a.hi, b);
  double cross = a.lo * b;
  p.lo += cross;
  return fast_two_sum(p.hi, p.lo);
}

DETM_HD dd dd_neg(dd a)
{
  return {-a.hi, -a.lo};
}

// log(1 + t) for |t| <= 2^-7 (+ tiny), t given as double-double.
DETM_HD dd log1p_small(dd t)
{
  double th = t.hi;
  // Taylor tail c_k = (-1)^(k+1)/k, k = 13 .. 5, in double.
  double q = 1.0 / 13.0;
  q = q * th + (-1.0 / 12.0);
  q = q * th + (1.0 / 11.0);
  q = q * th + (-1.0 / 10.0);
  q = q * th + (1.0 / 9.0);
  q = q * th + (-1.0 / 8.0);
  q = q * th + (1.0 / 7.0);
  q = q * th + (-1.0 / 6.0);
  q = q * th + (1.0 / 5.0);
  dd w = dd_mul_d(t, q);
  w = dd_add(w, dd {-0.25, 0.0});
  w = dd_mul(t, w);
  w = dd_add(w, dd {DETM_THIRD_HI, DETM_THIRD_LO});
  w = dd_mul(t, w);
  w = dd_add(w, dd {-0.5, 0.0});
  w = dd_mul(t, w);
  w = dd_add(w, dd {1.0, 0.0});
  return dd_mul(t, w);
}

//! Natural logarithm.
DETM_HD double log(double x)
{
  if (x == 1.0)
    return 0.0;
  if (!(x > 0.0)) {
    if (x == 0.0)
      return -__builtin_inf();
    return __builtin_nan("");
  }
  uint64_t ix = as_bits(x);
  if (ix >= 0x7ff0000000000000ULL)
    return x; // +inf
  int64_t scale = 0;
  if (ix < 0x0010000000000000ULL) {
    // subnormal: scale into the normal range
    ix = as_bits(x * 0x1p54);
    scale = -54;
  }
  const uint64_t off = 0x3fe6a00000000000ULL;
  uint64_t tmp = ix - off;
  int i = static_cast<int>((tmp >> 45) & 127);
  int64_t e = (static_cast<int64_t>(tmp) >> 52) + scale;
  double z = from_bits(ix - (tmp & 0xfff0000000000000ULL));
  double r = DETM_LOG_INVC[i];

  // t = z*r - 1 exactly, as a double-double: p = fl(z*r) lies in [0.5, 2] so
  // p - 1 is exact, and fma recovers the rounding error of the product.
  double p = z * r;
  double pe = fma_(z, r, -p);
  double th = p - 1.0;
  dd t = two_sum(th, pe);

  dd L = log1p_small(t);
  dd T = {DETM_LOG_NEGLOGR[2 * i], DETM_LOG_NEGLOGR[2 * i + 1]};
  double ed = static_cast<double>(e);
  dd A = fast_two_sum(ed * DETM_LN2_HI, ed * DETM_LN2_LO);
  dd s = dd_add(A, T);
  s = dd_add(s, L);
  return s.hi;
}

// sin and cos of y (|y| <= pi/4 + tiny), y as double-double.
DETM_HD void sincos_core(dd y, dd* s_out, dd* c_out)
{
  double jd = __builtin_floor(y.hi * 64.0 + 0.5);
  double a = jd * (1.0 / 64.0);
  dd z = two_sum(y.hi, -a);
  z.lo += y.lo;
  z = fast_two_sum(z.hi, z.lo);
  int j = static_cast<int>(jd);
  int ja = j < 0 ? -j : j;
  dd Sa = {DETM_SINCOS_TAB[4 * ja], DETM_SINCOS_TAB[4 * ja + 1]};
  dd Ca = {DETM_SINCOS_TAB[4 * ja + 2], DETM_SINCOS_TAB[4 * ja + 3]};
  if (j < 0)
    Sa = dd_neg(Sa);

  dd z2 = dd_mul(z, z);
  double h = z2.hi;
  // sin(z) = z + z * z^2 * (-1/6 + z^2/120 - z^4/5040 + z^6/362880)
  double ts = h * (1.0 / 120.0 + h * (-1.0 / 5040.0 + h * (1.0 / 362880.0)));
  dd ps = dd_add(dd {-DETM_SIXTH_HI, -DETM_SIXTH_LO}, dd {ts, 0.0});
  ps = dd_mul(z2, ps);
  dd sz = dd_add(z, dd_mul(z, ps));
  // cos(z) = 1 + z^2 * (-1/2 + z^2/24 - z^4/720 + z^6/40320)
  double tc = h * (1.0 / 24.0 + h * (-1.0 / 720.0 + h * (1.0 / 40320.0)));
  dd pc = dd_add(dd {-0.5, 0.0}, dd {tc, 0.0});
  pc = dd_mul(z2, pc);
  dd cz = dd_add(dd {1.0, 0.0}, pc);

  *s_out = dd_add(dd_mul(Sa, cz), dd_mul(Ca, sz));
  *c_out = dd_add(dd_mul(Ca, cz), dd_neg(dd_mul(Sa, sz)));
}

// Reduce x by multiples of pi/2: x = q*pi/2 + y.  Accurate for |x| < 2^19.
DETM_HD void reduce_pio2(double x, dd* y, int* q)
{
  double kd = __builtin_floor(x * DETM_TWO_OVER_PI + 0.5);
  double r1 = x - kd * DETM_PIO2_1; // exact (Sterbenz) for |kd| < 2^20
  double k2 = kd * DETM_PIO2_2;     // exact: 33-bit constant times |kd| < 2^20
  dd s = two_sum(r1, -k2);
  double k3 = kd * DETM_PIO2_3;
  s.lo -= k3;
  *y = fast_two_sum(s.hi, s.lo);
  *q = static_cast<int>(static_cast<int64_t>(kd) & 3);
}

//! sin(x) and cos(x) of the same argument; each equals detm::sin / detm::cos.
DETM_HD void sincos(double x, double* sx, double* cx)
{
  if (!(x - x == 0.0)) {
    *sx = __builtin_nan("");
    *cx = __builtin_nan("");
    return;
  }
  dd y;
  int q;
  reduce_pio2(x, &y, &q);
  dd s, c;
  sincos_core(y, &s, &c);
  switch (q) {
  case 0:
    *sx = s.hi;
    *cx = c.hi;
    break;
  case 1:
    *sx = c.hi;
    *cx = -s.hi;
    break;
  case 2:
    *sx = -s.hi;
    *cx = -c.hi;
    break;
  default:
    *sx = -c.hi;
    *cx = s.hi;
    break;
  }
}

DETM_HD double sin(double x)
{
  double s, c;
  sincos(x, &s, &c);
  return s;
}

DETM_HD double cos(double x)
{
  double s, c;
  sincos(x, &s, &c);
  return c;
}

//! Round half away from zero, as std::lround.
DETM_HD long lround(double x)
{
  double t = __builtin_trunc(x);
  double d = x - t; // exact
  if (d >= 0.5)
    t += 1.0;
  else if (d <= -0.5)
    t -= 1.0;
  return static_cast<long>(t);
}

} // namespace detm
