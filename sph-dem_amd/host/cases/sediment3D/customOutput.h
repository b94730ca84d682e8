// customOutput.h -- the sediment3D output plug-in (sediment3D/customOutput.h,
// customOutput.cpp): at each output step one line per DEM particle (time,
// vhat, fp, fv, fb, ff, porosity, r, fvel) into demVel.dat.
#pragma once

#include <fstream>

#include "customOutputBase.h"
#include "customSim.h"
#include "sph.h"

inline std::ofstream& demVelFile() {
  static std::ofstream fo;
  return fo;
}
inline void printVel(Cparticle& p, CglobalVars& g) {
  demVelFile() << g.time << ' ' << p.vhat << ' ' << p.fp << ' ' << p.fv << ' ' << p.fb << ' '
               << p.ff << ' ' << p.porosity << ' ' << p.r << ' ' << p.fvel << '\n';
}

class CcustomOutput : public CcustomOutputBase {
 public:
  CcustomOutput(CdataLL* _data) : CcustomOutputBase(_data) {
    demVelFile().open("demVel.dat");
    demVelFile() << "time vx vy vz fp_x fp_y fp_z fv_x fv_y fv_z fb_x fb_y fb_z ff_x ff_y ff_z "
                    "porosity rx ry rz fvelx fvely fvelz"
                 << std::endl;
  }
  void calcOutput(int outstep, CcustomSim* custSim, Cio_data_vtk* io) override {
    (void)outstep;
    (void)custSim;
    (void)io;
    data->traverse<printVel, Nsph::ifDem>();
  }
};
