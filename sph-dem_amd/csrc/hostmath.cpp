// Host build of sphdem_math.h: exposes the device arithmetic helpers so the
// CPU tests can check them against glibc and against each other.
#include "sphdem_math.h"

extern "C" {
double sphdem_host_pow_third(double x) { return sphdem::pow_third(x); }
double sphdem_host_wendW(double wcon, double q, double h) { return sphdem::wendW(wcon, q, h); }
double sphdem_host_wendF(double wcon, double q, double h) { return sphdem::wendF(wcon, q, h); }
double sphdem_host_wendPdWh(double wcon, double q, double h) { return sphdem::wendPdWh(wcon, q, h); }

// Array forms for bulk checks: out[i] = f(a[i], b[i]).
void sphdem_host_divr_arr(const double* x, const double* y, double* out, long n) {
  for (long i = 0; i < n; i++) out[i] = sphdem::divr(x[i], y[i], 1.0 / y[i]);
}
// the hoisted Wendland forms, with their per-particle reciprocals computed here
void sphdem_host_wendFi_arr(double wcon, const double* q, const double* h, double* out, long n) {
  for (long i = 0; i < n; i++) out[i] = sphdem::wendFi(wcon, q[i], 1.0 / sphdem::pow5(h[i]));
}
void sphdem_host_wendF_arr(double wcon, const double* q, const double* h, double* out, long n) {
  for (long i = 0; i < n; i++) out[i] = sphdem::wendF(wcon, q[i], h[i]);
}
void sphdem_host_wendWi_arr(double wcon, const double* q, const double* h, double* out, long n) {
  for (long i = 0; i < n; i++) out[i] = sphdem::wendWi(wcon, q[i], 1.0 / sphdem::cube(h[i]));
}
void sphdem_host_wendW_arr(double wcon, const double* q, const double* h, double* out, long n) {
  for (long i = 0; i < n; i++) out[i] = sphdem::wendW(wcon, q[i], h[i]);
}
void sphdem_host_wendPdWhi_arr(double wcon, const double* q, const double* h, double* out, long n) {
  for (long i = 0; i < n; i++) {
    const double hh = h[i];
    out[i] = sphdem::wendPdWhi(wcon, q[i], hh, 1.0 / hh, 1.0 / sphdem::cube(hh),
                               3.0 / sphdem::pow4(hh));
  }
}
void sphdem_host_wendPdWh_arr(double wcon, const double* q, const double* h, double* out, long n) {
  for (long i = 0; i < n; i++) out[i] = sphdem::wendPdWh(wcon, q[i], h[i]);
}
}
