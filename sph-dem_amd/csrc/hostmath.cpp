// Host build of sphdem_math.h: exposes the device arithmetic helpers so the
// CPU tests can check them against glibc.
#include "sphdem_math.h"

extern "C" {
double sphdem_host_pow_third(double x) { return sphdem::pow_third(x); }
double sphdem_host_wendW(double wcon, double q, double h) { return sphdem::wendW(wcon, q, h); }
double sphdem_host_wendF(double wcon, double q, double h) { return sphdem::wendF(wcon, q, h); }
double sphdem_host_wendPdWh(double wcon, double q, double h) { return sphdem::wendPdWh(wcon, q, h); }
}
