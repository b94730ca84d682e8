// sphdem_math.h -- arithmetic shared by the gfx950 kernels and their host-side
// checks.  Every expression keeps the operand order of the reference so that,
// compiled without FMA contraction, the device reproduces its rounding:
//   dot(a,b) = a0*b0 + (a1*b1 + a2*b2)   blitz/meta/dot.h:40-60
//   Wendland F, K, pdKh, W, pdWh          sph.h:74-81, 199-205, 285-292, 306-312
#pragma once

#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define SD_HD __host__ __device__ __forceinline__
#else
#define SD_HD static inline
#endif

namespace sphdem {

SD_HD double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  return a0 * b0 + (a1 * b1 + a2 * b2);
}

// Small integer powers, correctly rounded (up to rare near-midpoint cases)
// like libm's pow(x, n) that the reference calls: the exact power is carried
// as a double-double through FMA products and rounded once.  pow(x,2) is the
// plain product, which is correctly rounded by itself.
SD_HD double sq(double x) { return x * x; }
SD_HD double cube(double x) {
  const double p = x * x, e = fma(x, x, -p);      // x^2 = p + e exactly
  const double hi = p * x;
  const double lo = fma(p, x, -hi) + e * x;        // x^3 - hi
  return hi + lo;
}
SD_HD double pow4(double x) {
  const double p = x * x, e = fma(x, x, -p);
  const double hi = p * p;
  const double lo = fma(p, p, -hi) + 2.0 * (p * e);  // (p+e)^2 - hi, e^2 dropped
  return hi + lo;
}
SD_HD double pow5(double x) {
  const double p = x * x, e = fma(x, x, -p);
  const double q = p * p, eq = fma(p, p, -q) + 2.0 * (p * e);  // x^4 = q + eq
  const double hi = q * x;
  const double lo = fma(q, x, -hi) + eq * x;
  return hi + lo;
}

// Wendland C2 in 3D with support q <= 2 (KERNAL_RADIUS = 2).
SD_HD double wendF(double wcon, double q, double h) {
  if (q <= 2.0) {
    double t = 2.0 - q;
    return (1.0 / pow5(h)) * wcon * (-4.0 * cube(t) * (1.0 + 2.0 * q) + 2.0 * pow4(t)) / q;
  }
  return 0.0;
}
SD_HD double wendK(double wcon, double q) {
  if (q <= 2.0) return wcon * pow4(2.0 - q) * (1.0 + 2.0 * q);
  return 0.0;
}
SD_HD double wendPdKh(double wcon, double q, double h) {
  if (q <= 2.0) {
    double t = 2.0 - q;
    return -wcon * q * (-4.0 * cube(t) * (1.0 + 2.0 * q) + 2.0 * pow4(t)) / h;
  }
  return 0.0;
}
SD_HD double wendW(double wcon, double q, double h) { return (1.0 / cube(h)) * wendK(wcon, q); }
SD_HD double wendPdWh(double wcon, double q, double h) {
  return (1.0 / cube(h)) * wendPdKh(wcon, q, h) - 3.0 / pow4(h) * wendK(wcon, q);
}

// x / y correctly rounded, given ry = x-independent 1/y (itself correctly
// rounded, i.e. computed once with a real division).  Markstein's theorem:
// q0 = RN(x * ry) lies within one ulp of x/y, the remainder x - q0*y is exact
// in one FMA, and RN(q0 + rem*ry) is then RN(x/y).  Three FMA-class operations
// in place of a full division wherever one divisor serves many quotients.
// Valid for finite operands without overflow or subnormal quotients (the only
// difference otherwise: x = -0 gives +0).
SD_HD double divr(double x, double y, double ry) {
  const double q0 = x * ry;
  const double rem = fma(-q0, y, x);
  return fma(rem, ry, q0);
}

// The Wendland factors with the h-only parts computed once per particle:
//   i5 = 1/pow5(h), rh = 1/h, i3 = 1/cube(h), c3 = 3/pow4(h).
// Each equals the corresponding function above bit for bit.
SD_HD double wendPoly(double q) {
  const double t = 2.0 - q;
  return -4.0 * cube(t) * (1.0 + 2.0 * q) + 2.0 * pow4(t);
}
SD_HD double wendFi(double wcon, double q, double i5) {
  if (q <= 2.0) return i5 * wcon * wendPoly(q) / q;
  return 0.0;
}
SD_HD double wendPdKhi(double wcon, double q, double h, double rh) {
  if (q <= 2.0) return divr(-wcon * q * wendPoly(q), h, rh);
  return 0.0;
}
SD_HD double wendWi(double wcon, double q, double i3) { return i3 * wendK(wcon, q); }
SD_HD double wendPdWhi(double wcon, double q, double h, double rh, double i3, double c3) {
  return i3 * wendPdKhi(wcon, q, h, rh) - c3 * wendK(wcon, q);
}

// Correctly rounded pow(x, y) for y = the double nearest 1/3, x > 0 finite.
// y = 1/3 - d with d = 1/(3*2^54) exactly, so pow(x,y) = cbrt(x) * x^(-d) and
// x^(-d) = 1 - d ln x to far below an ulp.  cbrt(x) is carried as a double-
// double (one Newton step on c^3 - x evaluated with exact products), the
// product with (1 - d ln x) is added to the low part, and the sum is rounded
// once.  glibc's pow is correctly rounded on these arguments, which is what
// makes h = HFAC*pow(m/rho, 1.0/NDIM) (sphIncompress.h:699) identical on both
// sides; tests/test_math.py checks the two against each other.
SD_HD double two_prod_err(double a, double b, double p) { return fma(a, b, -p); }

SD_HD double pow_third(double x) {
  double c = cbrt(x);
  // c^2 = p2 + e2 exactly
  double p2 = c * c;
  double e2 = two_prod_err(c, c, p2);
  // c^3 = c*(p2 + e2) = p3 + e3 (+ c*e2 rounded, ~2^-106 relative)
  double p3 = c * p2;
  double e3 = two_prod_err(c, p2, p3) + c * e2;
  // residual c^3 - x, evaluated so that the large parts cancel exactly
  double res = (p3 - x) + e3;
  double lo = -res / (3.0 * p2);  // Newton correction: cbrt(x) = c + lo
  const double d = 1.0 / (3.0 * 18014398509481984.0);  // 1/(3*2^54)
  double corr = lo - (c * d) * log(x);
  return c + corr;
}

}  // namespace sphdem
