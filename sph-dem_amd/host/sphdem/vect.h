// vect.h -- the 3-vector of the particle record (the reference's blitz
// TinyVector<double,3>, vect.h:7-40) with the operations host functors use.
// dot() keeps blitz's association a0*b0 + (a1*b1 + a2*b2)
// (blitz/meta/dot.h:40-60); len2/len are vect.cpp:3-9.
#pragma once

#include <cmath>

namespace sphdem_host {

struct vect {
  double c[3];
  vect() : c{0.0, 0.0, 0.0} {}
  vect(double x, double y, double z) : c{x, y, z} {}
  vect& operator=(double s) {  // blitz: v = 0.0 sets every component
    c[0] = c[1] = c[2] = s;
    return *this;
  }
  double& operator[](int i) { return c[i]; }
  double operator[](int i) const { return c[i]; }
  vect& operator+=(const vect& o) {
    for (int i = 0; i < 3; i++) c[i] += o.c[i];
    return *this;
  }
  vect& operator-=(const vect& o) {
    for (int i = 0; i < 3; i++) c[i] -= o.c[i];
    return *this;
  }
  vect& operator*=(double s) {
    for (int i = 0; i < 3; i++) c[i] *= s;
    return *this;
  }
};
inline vect operator+(vect a, const vect& b) { return a += b; }
inline vect operator-(vect a, const vect& b) { return a -= b; }
inline vect operator*(vect a, double s) { return a *= s; }
inline vect operator*(double s, vect a) {
  for (int i = 0; i < 3; i++) a.c[i] = s * a.c[i];
  return a;
}
inline vect operator/(vect a, double s) {
  for (int i = 0; i < 3; i++) a.c[i] /= s;
  return a;
}
inline double dot(const vect& a, const vect& b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }
inline double len2(const vect& a) { return dot(a, a); }
inline double len(const vect& a) { return std::sqrt(len2(a)); }

}  // namespace sphdem_host
