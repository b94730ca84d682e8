// sphIncompress.h -- CsphIncompress(CdataLL*) with start(), middle(), end()
// (sphIncompress.cpp:3-251) and calcOutputVars() (sphIncompress.h:271-319),
// each one call into the library.  Host writes reach the device before each
// call and the host globals are written before and read back after it, so a
// CcustomSim sees the time and custom[] values the reference would show it
// between the phases.
#pragma once

#include "dataLL.impl.h"
#include "sph.h"

namespace sphdem_host {

class CsphIncompress {
 public:
  explicit CsphIncompress(CdataLL* _data) : data(_data) {}
  void start() { phase(sphdem_step_start); }
  void middle() { phase(sphdem_step_middle); }
  void end() { phase(sphdem_step_end); }
  // the output step's diagnostics into data->globals (momenta, energies,
  // maxima, density statistics, counts)
  void calcOutputVars() {
    data->flush();
    sphdem_output o;
    data->check(sphdem_output_vars(data->context(), &o));
    CglobalVars& g = data->globals;
    g.linMom = vect(o.linMom[0], o.linMom[1], o.linMom[2]);
    g.angMom = o.angMom;
    g.eKE = o.eKE;
    g.eViscF = o.eViscF;
    g.edViscFdt = o.edViscFdt;
    g.eViscB = o.eViscB;
    g.eBForce = o.eBForce;
    g.eFF = o.eFF;
    g.edFFdt = o.edFFdt;
    g.eElast = o.eElast;
    g.edElastdt = o.edElastdt;
    g.eElastExact = o.eElastExact;
    g.eTotal = o.eTotal;
    g.maxV = o.maxV;
    g.maxF = o.maxF;
    g.maxFF = o.maxFF;
    g.rmsFF = o.rmsFF;
    g.aveDens = o.aveDens;
    g.varDens = o.varDens;
    g.aveDensFromMass = o.aveDensFromMass;
    g.n = (int)o.n;
    g.nSph = (int)o.nSph;
  }

 private:
  void phase(int (*fn)(sphdem_ctx*)) {
    data->flush();
    data->check(fn(data->context()));
    data->deviceChanged();
  }
  CdataLL* data;
};

}  // namespace sphdem_host

using sphdem_host::CsphIncompress;
using namespace Nsph;  // as the reference's sphIncompress.h does (:13)
