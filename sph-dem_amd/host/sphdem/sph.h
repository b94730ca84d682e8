// sph.h -- what host case code reads of the reference's Nsph operator
// namespace (sph.h:27-42, 74-342, 1495-1516, 1822-1889): the Wendland C2
// kernel family (3D, support 2h; the same expressions, so host-side sums
// round as the reference's do), the linear spring-dashpot normal force, the
// particle-type predicates, h from density, and the global constants
// WCON_WENDLAND and ALPHA.  The timestep's own operators run on the GPU; these
// are for the functors a case applies between the half-steps.
#pragma once

#include <cmath>

#include "customConstants.h"
#include "demConstants.h"
#include "globalVars.h"
#include "misc.h"
#include "particle.h"
#include "vect.h"

#ifndef NDIM
#define NDIM 3
#endif

#ifndef SPHDEM_DERIVED_CONSTANTS
const double WCON_WENDLAND = 21.0 / (256.0 * PI);
#endif
const double ALPHA = Nmisc::viscToAlpha(VISCOSITY, H, SPSOUND);

namespace Nsph {

// the gradient factor: gradW = dr F(q, h)
inline double F(const double q, const double h) {
  if (q <= 2.0)
    return (1 / std::pow(h, NDIM + 2)) * WCON_WENDLAND *
           (-4 * std::pow(2 - q, 3) * (1 + 2 * q) + 2 * std::pow(2 - q, 4)) / q;
  return 0.0;
}
inline double K(const double q, const double h) {
  (void)h;
  if (q <= 2.0) return WCON_WENDLAND * std::pow(2.0 - q, 4) * (1.0 + 2.0 * q);
  return 0.0;
}
inline double pdKh(const double q, const double h) {
  if (q <= 2.0)
    return -WCON_WENDLAND * q * (-4 * std::pow(2 - q, 3) * (1 + 2 * q) + 2 * std::pow(2 - q, 4)) /
           h;
  return 0.0;
}
inline double W(const double q, const double h) { return (1 / std::pow(h, NDIM)) * K(q, h); }
inline double pdWh(const double q, const double h) {
  return (1 / std::pow(h, NDIM)) * pdKh(q, h) - NDIM / std::pow(h, NDIM + 1) * K(q, h);
}

#ifdef LIQ_DEM
// demNormal, LINEAR (sph.h:1495-1516)
inline vect demNormal(const double overlap, const double overlap_dot, const vect normal) {
  return (DEM_K * overlap + DEM_GAMMA * overlap_dot) * normal;
}
#endif

inline void calcHFromDens(Cparticle& p, CglobalVars&) {
  p.h = HFAC * std::pow(p.mass / p.dens, 1.0 / NDIM);
}

inline bool ifSph(Cparticle& p) { return p.iam == sph; }
inline bool ifDem(Cparticle& p) { return p.iam == dem; }
inline bool ifDemOrGhost(Cparticle& p) { return p.iam == dem || p.iam == ghost; }
inline bool ifDemOrDemBoundary(Cparticle& p) { return p.iam == dem; }
inline bool ifSphOrDem(Cparticle& p) { return p.iam == sph || p.iam == dem; }
inline bool ifSphOrBoundary(Cparticle& p) { return p.iam == sph || p.iam == boundary; }
inline bool ifSphOrSphBoundaryOrBoundary(Cparticle& p) {
  return p.iam == sph || p.iam == boundary || p.iam == sphBoundary;
}
inline bool ifSphBoundary(Cparticle& p) { return p.iam == sphBoundary; }
inline bool ifSphOrSphBoundary(Cparticle& p) { return p.iam == sph || p.iam == sphBoundary; }
// (the reference's "...OrImmersedDem" predicates test sph and sphBoundary only)
inline bool ifSphOrSphBoundaryOrImmersedDem(Cparticle& p) {
  return p.iam == sph || p.iam == sphBoundary;
}
inline bool ifSphOrSphBoundaryOrDem(Cparticle& p) {
  return p.iam == sph || p.iam == sphBoundary || p.iam == dem;
}
inline bool ifSphOrSphBoundaryOrImmersedDemOrDem(Cparticle& p) {
  return p.iam == sph || p.iam == sphBoundary || p.iam == dem;
}
inline bool ifSphOrSphBoundaryOrGhost(Cparticle& p) {
  return p.iam == sph || p.iam == sphBoundary || p.iam == ghost;
}
inline bool ifSphOrSphBoundaryOrDemOrGhost(Cparticle& p) {
  return p.iam == sph || p.iam == sphBoundary || p.iam == dem || p.iam == ghost;
}
inline bool ifSphOrGhost(Cparticle& p) { return p.iam == sph || p.iam == ghost; }
inline bool ifBoundary(Cparticle& p) {
  return p.iam == boundary || (p.iam >= boundaryobj1 && p.iam <= boundaryobj6);
}
inline bool ifBoundaryOrSphBoundary(Cparticle& p) {
  return p.iam == boundary || p.iam == sphBoundary;
}

}  // namespace Nsph
