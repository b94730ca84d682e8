// particle.h -- the particle record host functors see (Cparticle, the fields
// of particle.h:62-316 that the timestep loop touches) and the particle
// types (enum iamTypes, particle.h:17).  The container is indexed by the
// order the particles were handed to CdataLL; tag = index + 1.
#pragma once

#include <vector>

#include "vect.h"

namespace sphdem_host {

enum iamTypes {
  sph = 0,
  sphBoundary = 1,
  immovable = 2,
  boundary = 3,
  ghost = 4,
  dem = 11,
  deleted = 255,  // no counterpart: a particle the library removed as out of range
};

struct Cparticle {
  vect r, v, v0, vhat, vhat0, f, fp, fv, fb, fdrag, fvel;
  double dens = 0.0, dddt = 0.0, h = 0.0, mass = 0.0, press = 0.0, pdr2 = 0.0, omega = 0.0,
         porosity = 1.0, shepSum = 0.0, u = 0.0, dudt = 0.0;
  int iam = sph;
  int tag = 0;
};

typedef std::vector<Cparticle> particleContainer;

}  // namespace sphdem_host
