// globalVars.h -- the CglobalVars members the timestep loop, the output step
// and a one-rank run.cpp read and write (globalVars.h:11-235 of the
// reference): time, dt, newDt, maxdt, sphStep, custom[], the rank's domain
// procDomain[], the output-step diagnostics calcOutputVars fills, and the
// single-rank MPI fields (mpiRank 0 of mpiSize 1).
#pragma once

#include "vect.h"

namespace sphdem_host {

struct CglobalVars {
  double time = 0.0, dt = 0.0, newDt = 0.0, maxdt = 0.0;
  long long sphStep = 0;
  int outstep = 0;
  double custom[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double procDomain[6] = {0, 0, 0, 0, 0, 0};
  int mpiRank = 0, mpiSize = 1;
  double tmp = 0.0;
  // calcOutputVars (sphIncompress.h:271-319)
  vect linMom;
  double angMom = 0.0, eKE = 0.0, eViscF = 0.0, edViscFdt = 0.0, eViscB = 0.0, eBForce = 0.0,
         eFF = 0.0, edFFdt = 0.0, eElast = 0.0, edElastdt = 0.0, eElastExact = 0.0,
         eTotal = 0.0, maxV = 0.0, maxF = 0.0, maxFF = 0.0, rmsFF = 0.0, aveDens = 0.0,
         varDens = 0.0, aveDensFromMass = 0.0;
  int n = 0, nSph = 0;
  // the per-output-step wall-clock bookkeeping run.cpp calls (globalVars.h:163-188);
  // the library times its own kernels (sphdem_timing_begin/end)
  void startOutStep() {}
  void endOutStep() {}
};

}  // namespace sphdem_host

using sphdem_host::CglobalVars;
