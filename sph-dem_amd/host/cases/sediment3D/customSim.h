// customSim.h -- the sediment3D case plug-in (sediment3D/customSim.cpp:5-113)
// written against the C++ host: host functors applied with data->traverse,
// exactly the hooks the reference case runs around the half-steps.
#pragma once

#include "../../sphdem/customSimBase.h"
#include "customConstants.h"

namespace sphdem_host {

inline bool ifSph(Cparticle& p) { return p.iam == sph; }
inline bool ifDem(Cparticle& p) { return p.iam == dem; }
inline bool ifSphOrDem(Cparticle& p) { return p.iam == sph || p.iam == dem; }

// sediment3D/customSim.cpp:5-7
inline void addGravity(Cparticle& p, CglobalVars&) { p.f[2] -= 9.81; }
// sediment3D/customSim.cpp:12-25
inline void freezeParticles(Cparticle& p, CglobalVars&) {
  p.f = 0.0;
  p.v = 0.0;
  p.vhat = p.v;
}
// sediment3D/customSim.cpp:27-42
inline void sumDragForces(Cparticle& p, CglobalVars& g) {
  if (p.iam == sph) {
    for (int i = 0; i < 3; i++) g.custom[i] += p.mass * p.fdrag[i];
  } else if (p.iam == dem) {
    for (int i = 3; i < 6; i++) g.custom[i] += p.mass * p.fdrag[i - 3];
  }
}
// sediment3D/customSim.cpp:67-81, the linear normal force of Nsph::demNormal
// (sph.h:1676-1690) against the floor z = RMIN[2]
inline void addCustomBoundaries(Cparticle& p, CglobalVars&) {
  const double overlap = RMIN[2] + DEM_RADIUS - p.r[2];
  if (overlap > 0.0) {
    const double overlap_dot = -p.vhat[2];
    const double s = DEM_K * overlap + DEM_GAMMA * overlap_dot;
    const vect fNorm(0.0 * s / p.mass, 0.0 * s / p.mass, 1.0 * s / p.mass);
    p.fb += fNorm;
    p.f += fNorm;
  }
}

class CcustomSim : public CcustomSimBase {
 public:
  CcustomSim(CdataLL* _data, double _time) : CcustomSimBase(_data, _time) {}
  void beforeMiddle(double newTime) override {
    time = newTime;
    data->traverse<addGravity, ifSph>();
  }
  void beforeEnd(double newTime) override {
    time = newTime;
    for (int i = 0; i < 6; i++) data->globals.custom[i] = 0.0;
    data->traverse<sumDragForces, ifSphOrDem>();
    data->traverse<addGravity, ifDem>();
    data->traverse<addCustomBoundaries, ifDem>();
    if (data->globals.time <= TIME_DROP_PARTICLE) data->traverse<freezeParticles, ifDem>();
  }
  void afterEnd(double newTime) override {
    time = newTime;
    if (data->globals.time <= TIME_DROP_PARTICLE) data->traverse<freezeParticles, ifDem>();
  }
};

}  // namespace sphdem_host
