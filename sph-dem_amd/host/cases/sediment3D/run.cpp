// run.cpp -- the sediment3D case driven as the reference's run.cpp:73-160
// drives a case on one rank: CdataLL over the particle container,
// CsphIncompress, CcustomSim and CcustomOutput, then per step sphStep++,
// beforeStart, start, beforeMiddle, middle, beforeEnd, end, afterEnd, and at
// the end an output step (calcOutputVars, calcOutput, the output, globals and
// restart files, calcTraverseOrder).
//
// usage: run <initial condition file> <steps> <state file> [<output base name>]
//   initial condition: magic "SPHDEMIC", int64 n, r v vhat (3n doubles),
//   mass h dens (n doubles), iam (n int32)
//   state (written): magic "SPHDEMST", int64 n, time, dt, then r v vhat f
//   (3n), dens h press omega porosity (n), per particle in input order
//   output base name: if given, the output step writes <base>0000001.pvtu
//   (+ _0.vtu), <base>Globals0.dat and <base>Restart0000001.pvtu (+ _0.vtu)
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "customConstants.h"
#include "customOutput.h"
#include "customSim.h"
#include "io_data_vtk.h"
#include "sphIncompress.h"

static bool readAll(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

int main(int argc, char** argv) {
  if (argc != 4 && argc != 5) {
    fprintf(stderr, "usage: %s <initial condition> <steps> <state out> [<output base>]\n",
            argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  char magic[8];
  int64_t n = 0;
  if (!f || !readAll(f, magic, 8) || memcmp(magic, "SPHDEMIC", 8) != 0 || !readAll(f, &n, 8) ||
      n <= 0) {
    fprintf(stderr, "run: %s is not an initial condition file\n", argv[1]);
    return 2;
  }
  std::vector<double> r(3 * n), v(3 * n), vh(3 * n), m(n), h(n), d(n);
  std::vector<int32_t> iam(n);
  bool ok = readAll(f, r.data(), 24 * n) && readAll(f, v.data(), 24 * n) &&
            readAll(f, vh.data(), 24 * n) && readAll(f, m.data(), 8 * n) &&
            readAll(f, h.data(), 8 * n) && readAll(f, d.data(), 8 * n) &&
            readAll(f, iam.data(), 4 * n);
  fclose(f);
  if (!ok) {
    fprintf(stderr, "run: %s is truncated\n", argv[1]);
    return 2;
  }
  CglobalVars globals;
  particleContainer ps(n);
  for (int64_t i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) {
      ps[i].r[k] = r[3 * i + k];
      ps[i].v[k] = v[3 * i + k];
      ps[i].vhat[k] = vh[3 * i + k];
    }
    ps[i].mass = m[i];
    ps[i].h = h[i];
    ps[i].dens = d[i];
    ps[i].iam = iam[i];
  }
  const sphdem_params P = sphdem_host::paramsFromConstants();
  for (int i = 0; i < 6; i++) globals.procDomain[i] = P.proc_domain[i];

  // run.cpp:77-87
  CdataLL* data = new CdataLL(ps, globals, true);
  CsphIncompress sph(data);
  CcustomSim customSim(data, globals.time);
  CcustomOutput customOutput(data);
  globals.maxdt = MAXDT;

  // run.cpp:102-119
  const int steps = atoi(argv[2]);
  for (int s = 0; s < steps; s++) {
    globals.sphStep++;
    customSim.beforeStart(globals.time);
    sph.start();
    customSim.beforeMiddle(globals.time);
    sph.middle();
    customSim.beforeEnd(globals.time);
    sph.end();
    customSim.afterEnd(globals.time);
  }

  // an output step (run.cpp:120-141)
  int outstep = 1;
  globals.outstep = outstep;
  globals.endOutStep();
  sph.calcOutputVars();
  if (argc == 5) {
    Cio_data_vtk io_data(argv[4], &globals);
    customOutput.calcOutput(outstep, &customSim, &io_data);
    io_data.writeOutput(outstep, *data->getParticles(), &globals);
    io_data.writeGlobals(outstep, &globals);
    io_data.writeRestart(outstep, *data->getParticles(), &globals);
    io_data.writeDomain(outstep, &globals);
  }
  data->calcTraverseOrder();
  globals.startOutStep();
  data->getParticles();  // the host view, current with the GPU

  FILE* o = fopen(argv[3], "wb");
  if (!o) {
    fprintf(stderr, "run: cannot write %s\n", argv[3]);
    return 2;
  }
  fwrite("SPHDEMST", 1, 8, o);
  fwrite(&n, 8, 1, o);
  fwrite(&globals.time, 8, 1, o);
  fwrite(&globals.dt, 8, 1, o);
  std::vector<double> buf(3 * n);
  vect Cparticle::*vecs[4] = {&Cparticle::r, &Cparticle::v, &Cparticle::vhat, &Cparticle::f};
  for (auto mv : vecs) {
    for (int64_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) buf[3 * i + k] = (ps[i].*mv)[k];
    fwrite(buf.data(), 8, 3 * n, o);
  }
  double Cparticle::*scl[5] = {&Cparticle::dens, &Cparticle::h, &Cparticle::press,
                               &Cparticle::omega, &Cparticle::porosity};
  for (auto ms : scl) {
    for (int64_t i = 0; i < n; i++) buf[i] = ps[i].*ms;
    fwrite(buf.data(), 8, n, o);
  }
  fclose(o);
  printf("run: %d steps, time %.9g, dt %.9g, eKE %.9g, field uploads %ld\n", steps, globals.time,
         globals.dt, globals.eKE, data->fieldUploads());
  delete data;
  return 0;
}
