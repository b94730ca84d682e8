// globalVars.h -- the CglobalVars members the timestep loop reads and writes
// (globalVars.h:11-235 of the reference): time, dt, newDt, maxdt, sphStep,
// custom[] and the rank's domain procDomain[].
#pragma once

namespace sphdem_host {

struct CglobalVars {
  double time = 0.0, dt = 0.0, newDt = 0.0, maxdt = 0.0;
  long long sphStep = 0;
  double custom[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double procDomain[6] = {0, 0, 0, 0, 0, 0};
};

}  // namespace sphdem_host
