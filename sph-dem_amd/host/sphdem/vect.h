// vect.h -- the 3-vector of the particle record (the reference's blitz
// TinyVector<double,3>, vect.h:7-40) with the operations host functors use,
// and vectInt (TinyVector<int,3>).  dot() keeps blitz's association
// a0*b0 + (a1*b1 + a2*b2) (blitz/meta/dot.h:40-60); len2/len are vect.cpp:3-9.
//
// Case code written for the reference names std members (vector, cout,
// ofstream, endl) unqualified, which its build got from blitz's headers; this
// header brings namespace std into the global namespace for the same reason,
// and exports the drop-in's names there.
#pragma once

#include <cmath>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace sphdem_host {

struct vect {
  double c[3];
  vect() : c{0.0, 0.0, 0.0} {}
  vect(double x, double y, double z) : c{x, y, z} {}
  vect(double s) : c{s, s, s} {}  // blitz: vect v = 0.0 fills every component
  vect& operator=(double s) {
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
  vect& operator/=(double s) {
    for (int i = 0; i < 3; i++) c[i] /= s;
    return *this;
  }
};
inline vect operator+(vect a, const vect& b) { return a += b; }
inline vect operator-(vect a, const vect& b) { return a -= b; }
inline vect operator-(vect a) {
  for (int i = 0; i < 3; i++) a.c[i] = -a.c[i];
  return a;
}
inline vect operator*(vect a, double s) { return a *= s; }
inline vect operator*(double s, vect a) {
  for (int i = 0; i < 3; i++) a.c[i] = s * a.c[i];
  return a;
}
inline vect operator/(vect a, double s) { return a /= s; }
inline bool operator!=(const vect& a, double s) {
  return a.c[0] != s || a.c[1] != s || a.c[2] != s;
}
inline double dot(const vect& a, const vect& b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }
inline double len2(const vect& a) { return dot(a, a); }
inline double len2(double a) { return a * a; }
inline double len(const vect& a) { return std::sqrt(len2(a)); }
inline vect cross(const vect& a, const vect& b) {
  return vect(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline std::ostream& operator<<(std::ostream& o, const vect& v) {
  return o << v[0] << ' ' << v[1] << ' ' << v[2];
}

struct vectInt {
  int c[3];
  vectInt() : c{0, 0, 0} {}
  vectInt(int x, int y, int z) : c{x, y, z} {}
  int& operator[](int i) { return c[i]; }
  int operator[](int i) const { return c[i]; }
};

}  // namespace sphdem_host

using namespace std;
using sphdem_host::cross;
using sphdem_host::dot;
using sphdem_host::len;
using sphdem_host::len2;
using sphdem_host::vect;
using sphdem_host::vectInt;
