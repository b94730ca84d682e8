// customSim.h -- the sediment3D case plug-in (sediment3D/customSim.cpp:5-113)
// written against the C++ host: host functors applied with data->traverse,
// exactly the hooks the reference case runs around the half-steps.
#pragma once

#include "customConstants.h"
#include "customSimBase.h"
#include "dataLL.h"
#include "sph.h"

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
// sediment3D/customSim.cpp:67-81: the floor z = RMIN[2] with Nsph::demNormal
inline void addCustomBoundaries(Cparticle& p, CglobalVars&) {
  const double overlap = RMIN[2] + DEM_RADIUS - p.r[2];
  if (overlap > 0.0) {
    const vect normal(0.0, 0.0, 1.0);
    const double overlap_dot = -p.vhat[2];
    const vect fNorm = Nsph::demNormal(overlap, overlap_dot, normal) / p.mass;
    p.fb += fNorm;
    p.f += fNorm;
  }
}

class CcustomSim : public CcustomSimBase {
 public:
  CcustomSim(CdataLL* _data, double _time) : CcustomSimBase(_data, _time) {}
  void beforeMiddle(double newTime) override {
    time = newTime;
    data->traverse<addGravity, Nsph::ifSph>();
  }
  void beforeEnd(double newTime) override {
    time = newTime;
    for (int i = 0; i < 6; i++) data->globals.custom[i] = 0.0;
    data->traverse<sumDragForces, Nsph::ifSphOrDem>();
    data->traverse<addGravity, Nsph::ifDem>();
    data->traverse<addCustomBoundaries, Nsph::ifDem>();
    if (data->globals.time <= TIME_DROP_PARTICLE) data->traverse<freezeParticles, Nsph::ifDem>();
  }
  void afterEnd(double newTime) override {
    time = newTime;
    if (data->globals.time <= TIME_DROP_PARTICLE) data->traverse<freezeParticles, Nsph::ifDem>();
  }
};
