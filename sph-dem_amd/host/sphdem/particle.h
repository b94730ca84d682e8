// particle.h -- the particle record host functors see (Cparticle: the fields
// of particle.h:62-316 that the timestep loop, the case plug-ins and the
// restart files touch) and the particle types (enum iamTypes, particle.h:17,
// with the same integer values).  The container is indexed by the order the
// particles were handed to CdataLL; tag = index + 1.
#pragma once

#include <vector>

#include "sphConstants.h"
#include "vect.h"

namespace sphdem_host {

enum iamTypes {
  sph = 0,
  sphBoundary,
  immovable,
  boundary,
  ghost,
  boundaryobj1,
  boundaryobj2,
  boundaryobj3,
  boundaryobj4,
  boundaryobj5,
  boundaryobj6,
  dem,
  demBoundary,
  immersedDem,
  deleted = 255,  // no counterpart: a particle the library removed as out of range
};

struct Cparticle {
  vect r, v, v0, vhat, vhat0, f, fp, fv, fb, ff, fdrag, fvel;
  vect vort, norm1, norm2, norm3;
  double dens = 0.0, dddt = 0.0, h = 0.0, mass = 0.0, press = 0.0, pdr2 = 0.0, omega = 0.0,
         porosity = 1.0, shepSum = 0.0, u = 0.0, dudt = 0.0;
  double eViscF = 0.0, deViscFdt = 0.0, eViscB = 0.0, eBForce = 0.0, eFF = 0.0, deFFdt = 0.0;
  double colour = 0.0, alpha = 0.0, tmp = 0.0;
  int iam = sph;
  int tag = 0;
};

typedef std::vector<Cparticle> particleContainer;

}  // namespace sphdem_host

using sphdem_host::Cparticle;
using sphdem_host::iamTypes;
using sphdem_host::particleContainer;
using sphdem_host::sph;
using sphdem_host::sphBoundary;
using sphdem_host::immovable;
using sphdem_host::boundary;
using sphdem_host::ghost;
using sphdem_host::boundaryobj1;
using sphdem_host::boundaryobj2;
using sphdem_host::boundaryobj3;
using sphdem_host::boundaryobj4;
using sphdem_host::boundaryobj5;
using sphdem_host::boundaryobj6;
using sphdem_host::dem;
using sphdem_host::demBoundary;
using sphdem_host::immersedDem;
