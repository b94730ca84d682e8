// run.cpp -- the sediment3D case driven like the reference's run.cpp:80-119:
// CdataLL over the particle container, CsphIncompress, CcustomSim, then per
// step sphStep++, beforeStart, start, beforeMiddle, middle, beforeEnd, end,
// afterEnd.
//
// usage: run <initial condition file> <steps> <state file>
//   initial condition: magic "SPHDEMIC", int64 n, r v vhat (3n doubles),
//   mass h dens (n doubles), iam (n int32)
//   state (written): magic "SPHDEMST", int64 n, time, dt, then r v vhat f
//   (3n), dens h press omega porosity (n), per particle in input order
#include <cstdio>
#include <cstring>
#include <vector>

#include "customConstants.h"  // first: constants.h and the core read its names
#include "../../sphdem/constants.h"
#include "../../sphdem/sphIncompress.h"
#include "customSim.h"

using namespace sphdem_host;

static bool readAll(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s <initial condition> <steps> <state out>\n", argv[0]);
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
  const sphdem_params P = paramsFromConstants();
  CglobalVars globals;
  globals.maxdt = P.maxdt;
  for (int i = 0; i < 6; i++) globals.procDomain[i] = P.proc_domain[i];

  CdataLL data(ps, globals, false, P);
  CsphIncompress sph(&data);
  CcustomSim customSim(&data, globals.time);
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
  data.pull();

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
  printf("run: %d steps, time %.9g, dt %.9g\n", steps, globals.time, globals.dt);
  return 0;
}
