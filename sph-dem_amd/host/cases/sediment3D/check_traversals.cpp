// check_traversals.cpp -- the pair traversals of the C++ drop-in over the
// sediment3D case: neighbours<F, If>(), neighboursGroup<F, If>() (with a write
// to the particles that must reach the GPU), neighboursGroupCoupling<F, If>()
// and traverse<T, F, If>(T&), with host functors written as a reference case
// would write them (dataLL.impl.h:16-281).
//
// usage: check_traversals <initial condition file>
// prints one line of key=value counts for the caller to compare with the
// library's own pair sets.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "customConstants.h"
#include "dataLL.h"
#include "sphIncompress.h"

using namespace sphdem_host;

static bool readAll(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

static long gPairs = 0, gGroups = 0, gGroupNb = 0, gCpl = 0;
static double gWorst = 0.0, gCplWorst = 0.0;

static double dist2(const Cparticle& a, const Cparticle& b) {
  double s = 0.0;
  for (int k = 0; k < 3; k++) s += (a.r[k] - b.r[k]) * (a.r[k] - b.r[k]);
  return s;
}
// the neighbour criterion of is_neighbr (particle.h:236-238): |dr| <= 2 max(h)
static double criterion(const Cparticle& a, const Cparticle& b) {
  const double lim = 2.0 * std::max(a.h, b.h);
  return dist2(a, b) / (lim * lim);
}

static bool ifSphGroup(Cparticle& p) { return p.iam == sph || p.iam == sphBoundary; }
static bool isDem(Cparticle& p) { return p.iam == dem; }
static bool ifAny(Cparticle&) { return true; }

static void countPair(Cparticle& a, Cparticle& b, CglobalVars&) {
  gPairs++;
  gWorst = std::max(gWorst, criterion(a, b));
}
static void countGroup(Cparticle& p, std::vector<Cparticle*>& nb, CglobalVars&) {
  gGroups++;
  gGroupNb += (long)nb.size();
  p.u = (double)nb.size();  // written back to the GPU by the traversal
}
static void countCoupling(Cparticle& p, std::vector<Cparticle*>& nb, CglobalVars&) {
  gCpl += (long)nb.size();
  for (Cparticle* q : nb) gCplWorst = std::max(gCplWorst, criterion(p, *q));
}
static void sumMass(Cparticle& p, CglobalVars&, double& acc) { acc += p.mass; }
// a scatter into the partners (ghost copies included), carried back to the
// origins by reverseSyncParticlesBetweenProcs
static void zeroTmp(Cparticle& p, CglobalVars&) { p.tmp = 0.0; }
static void scatterOne(Cparticle& p, std::vector<Cparticle*>& nb, CglobalVars&) {
  (void)p;
  for (Cparticle* q : nb) q->tmp += 1.0;
}
static double readTmp(Cparticle& p) { return p.tmp; }
static void sumTmp(Cparticle& p, double x) { p.tmp += x; }
static void kickZ(Cparticle& p, CglobalVars&) { p.f[2] += 1.0; }

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s <initial condition>\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  char magic[8];
  int64_t n = 0;
  if (!f || !readAll(f, magic, 8) || memcmp(magic, "SPHDEMIC", 8) != 0 || !readAll(f, &n, 8) ||
      n <= 0) {
    fprintf(stderr, "check_traversals: %s is not an initial condition file\n", argv[1]);
    return 2;
  }
  std::vector<double> r(3 * n), v(3 * n), vh(3 * n), m(n), h(n), d(n);
  std::vector<int32_t> iam(n);
  bool ok = readAll(f, r.data(), 24 * n) && readAll(f, v.data(), 24 * n) &&
            readAll(f, vh.data(), 24 * n) && readAll(f, m.data(), 8 * n) &&
            readAll(f, h.data(), 8 * n) && readAll(f, d.data(), 8 * n) &&
            readAll(f, iam.data(), 4 * n);
  fclose(f);
  if (!ok) return 2;
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
  CglobalVars globals;
  globals.maxdt = MAXDT;
  CdataLL data(ps, globals, false);
  data.reset();
  data.neighbours<countPair>();
  data.neighboursGroup<countGroup, ifSphGroup>();
  data.neighboursGroupCoupling<countCoupling, isDem>();
  double mass = 0.0;
  data.traverse<double, sumMass, ifAny>(mass);
  // the group sizes written by countGroup came back from the GPU
  data.pull();
  long written = 0;
  for (const Cparticle& p : ps)
    if (p.iam == sph || p.iam == sphBoundary) written += (long)p.u;
  // the scatter: each particle ends with the number of lists it appears in,
  // ghost appearances included once reverseSyncParticlesBetweenProcs has
  // carried them to their origins (printed as sum and index-weighted sum, for
  // the caller to compare with the library's pair set)
  data.traverse<zeroTmp, ifAny>();
  data.neighboursGroup<scatterOne, ifSphGroup>();
  data.reverseSyncParticlesBetweenProcs<double, readTmp, sumTmp>();
  long scatterSum = 0, scatterW = 0;
  for (size_t i = 0; i < ps.size(); i++) {
    scatterSum += (long)ps[i].tmp;
    scatterW += (long)(i + 1) * (long)ps[i].tmp;
  }
  // lazy sync: a traversal that writes f uploads that one field
  const long up0 = data.fieldUploads();
  data.traverse<kickZ, ifAny>();
  data.reset();
  const long fieldUploads = data.fieldUploads() - up0;
  // the output diagnostics through CsphIncompress
  CsphIncompress sphInc(&data);
  sphInc.calcOutputVars();
  // insertion, deletion and tags
  const int n0 = data.numParticles();
  int demTag = -1;
  for (const Cparticle& p : ps)
    if (p.iam == dem) demTag = p.tag;
  Cparticle* victim = data.getParticleFromTag(demTag);
  const bool tagOk = victim && victim->tag == demTag && victim->iam == dem;
  Cparticle copy = *victim;
  data.markForDeletion(*victim);
  data.deleteParticles();
  const int nAfterDelete = data.numParticles();
  copy.tag = n0 + 1;
  copy.r[2] += 0.5 * PSEP;
  data.insertNewParticle(copy);
  data.reset();
  int64_t onDevice = 0;
  sphdem_num_particles(data.context(), &onDevice);
  Cparticle* found = data.getParticleFromTag(n0 + 1);
  const bool insertedOk = found && found->iam == dem && found->r[2] == copy.r[2];
  printf("pairs=%ld worst=%.9f groups=%ld groupnb=%ld written=%ld coupling=%ld cplworst=%.9f "
         "mass=%.17g scattersum=%ld scatterw=%ld fielduploads=%ld nsph=%d outn=%d tagok=%d ndel=%d "
         "ndev=%lld insertok=%d\n",
         gPairs, gWorst, gGroups, gGroupNb, written, gCpl, gCplWorst, mass, scatterSum, scatterW,
         fieldUploads, globals.nSph, globals.n, tagOk ? 1 : 0, n0 - nAfterDelete,
         (long long)onDevice, insertedOk ? 1 : 0);
  return 0;
}
