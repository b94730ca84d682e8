// run_shaped.cpp -- a one-rank driver in the shape of the reference's
// run.cpp:15-160 (restart read, CdataLL / CsphIncompress / CcustomSim /
// CcustomOutput construction, the step loop with the plug-in hooks, the
// output step with calcOutputVars, calcOutput, the output, globals, restart
// and domain files, calcTraverseOrder), compiled by
// tests/test_host_dropin.py against the drop-in headers together with a case
// directory's customSim.cpp / customOutput.cpp, to show that case code
// written for the reference builds unchanged.
#include <cstdlib>
#include <string>

#include "customConstants.h"
#include "customOutput.h"
#include "customSim.h"
#include "globalVars.h"
#include "io_data_vtk.h"
#include "particle.h"
#include "sphIncompress.h"
#include "vect.h"

int main(int argc, char* argv[]) {
  if (argc != 4) {
    cout << "Usage: run infilename intimestep outfilename" << endl;
    return (-1);
  }
  CglobalVars globals;
  particleContainer ps;
  string filename = argv[1];
  Cio_data_vtk io_data(filename, &globals);
  io_data.readDomain(atoi(argv[2]), &globals);
  io_data.readGlobals(atoi(argv[2]), &globals);
  io_data.readRestart(atoi(argv[2]), &ps, &globals);
  int outstep = atoi(argv[2]);
  io_data.setFilename(argv[3], &globals);

  CdataLL* data = new CdataLL(ps, globals, true);
  CsphIncompress sph(data);
  CcustomSim customSim(data, globals.time);
  CcustomOutput customOutput(data);
  globals.maxdt = MAXTIME / OUTSTEP;
  double nextWriteTime = globals.time + MAXTIME / OUTSTEP;
  globals.startOutStep();
  while ((globals.time <= MAXTIME) && (outstep <= OUTSTEP)) {
    globals.sphStep++;
    customSim.beforeStart(globals.time);
    sph.start();
    customSim.beforeMiddle(globals.time);
    sph.middle();
    customSim.beforeEnd(globals.time);
    sph.end();
    customSim.afterEnd(globals.time);
    if (globals.time >= nextWriteTime) {
      outstep++;
      globals.outstep = outstep;
      globals.endOutStep();
      sph.calcOutputVars();
      customOutput.calcOutput(outstep, &customSim, &io_data);
      io_data.writeOutput(outstep, ps, &globals);
      io_data.writeGlobals(outstep, &globals);
      if (outstep % RESTART_EVERY_N_STEPS == 0) {
        io_data.writeRestart(outstep, ps, &globals);
        io_data.writeDomain(outstep, &globals);
      }
      data->calcTraverseOrder();
      nextWriteTime = globals.time + MAXTIME / OUTSTEP;
      globals.maxdt = MAXTIME / OUTSTEP;
      globals.startOutStep();
    }
  }
  delete data;
  return 0;
}
