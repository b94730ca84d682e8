/*
 * sphdem_oracle.c -- CPU FP64 restatement of SPH-DEM's per-timestep particle
 * loop, used ONLY as the checker of the GPU library: tests/, the smoke check
 * and the CPU-baseline leg of the benchmark load it.  Product code never links
 * or calls it.
 *
 * What it restates (every function cites the reference file:line it follows):
 *   - CdataLL: reset/updateDomain (periodic wrap, ghost images, deletion),
 *     calcDomainLimits, cell binning with getCellI/dem_getCellI, lazily cached
 *     neighbour lists, the coupling search with its asymmetric cell rule.
 *   - Nsph: Wendland kernel family, calcOmega, calcDensity, viscous (Morris and
 *     artificial) and pressure pair forces with the per-pair dt conditions,
 *     porosity, Shepard sum (as the reference's scatter + reverse sync),
 *     DEM->SPH drag, calcFPI + calcParticleDrag (Di Felice), calcDEMContacts.
 *   - CsphIncompress: start/middle/end and the integrator operators.
 *   - sediment3D CcustomSim hooks.
 * Arithmetic follows the reference expression by expression: blitz dot() is
 * a0*b0 + (a1*b1 + a2*b2), no FMA contraction (build with -ffp-contract=off),
 * glibc pow/exp/log10/sqrt.
 *
 * Compiled by oracle/Makefile into oracle/liboracle.so and driven through
 * ctypes (tests/oracle.py).  The reference itself does not build in this image
 * (vect.h:7 includes GSL headers, which are absent), so this restatement is
 * pinned by closed forms, invariants and brute force in tests/test_oracle.py
 * and tests/test_oracle_physics.py.
 *
 * Parity status of each function is listed in the table at the end of this
 * header comment in the form  function : how it is pinned.
 *   kernel F/K/W/pdKh/pdWh       : closed form, quadrature (int W dV), FD of K
 *   cells getCellI/dem_getCellI  : formula restated; brute-force bin check
 *   calcNeighbours               : brute force O(N^2) with the same predicate
 *   periodic ghosts/updateDomain : brute-force minimum image pair set
 *   calcOmega, calcDensity       : lattice sums in closed form
 *   calcPressForce               : -grad(p)/rho of a linear pressure field on
 *                                  a lattice; zero on a uniform lattice;
 *                                  antisymmetry / momentum conservation
 *   calcViscForce (Morris)       : nu lap(v) of a quadratic shear (plain and
 *                                  LIQ_DEM coefficients), zero for a linear
 *                                  shear; Cleary/Monaghan: conservation,
 *                                  dissipation, zero under rigid rotation
 *   calcDddtDudt                 : 3 alpha rho under uniform compression
 *                                  (1/omega_a included); zero under rigid
 *                                  translation and rotation
 *   calcPressSpsoundPdr2         : closed form; LIQ_DEM zero at REFD*eps and
 *                                  slope PRB*GAMMA/(REFD*eps)
 *   integrator, dt rule          : hand-computed single-particle trajectory
 *                                  (also fixes the first-step dt = 0)
 *   calcParticleDrag             : Di Felice closed form
 *   calcFPI                      : uniform-fluid closed forms of every
 *                                  Shepard-normalised output, the fv double
 *                                  count included, plus the Di Felice term
 *   calcDEMContacts              : two-sphere restitution closed form
 *   porosity, calcShepSum,
 *   calcDEMtoSPHDrag             : porosity/Shepard-sum consistency; Newton's
 *                                  third law of the drag pair
 *   gradW's omega under VAR_H     : pa's for both kernel terms of a pair,
 *                                  pinned by the invariant sum m_a omega_a
 *                                  f_a = 0 on a box where omega varies
 *   calcNeighboursCoupling        : known answers worked by hand from the
 *     (SPH side)                   reference's text (whole cells beyond Rc
 *                                  towards negative offsets, skipped cells
 *                                  within Rc at mixed-sign offsets, the
 *                                  strict < at plus == t^2); a closest-corner
 *                                  rule, swapped plus/minus, <= or a sign flip
 *                                  each fail one; plus brute force
 *   parity unpinned (only the compiled reference could pin it; restated
 *   from the cited lines): the order of the reverse-sync additions of the
 *   Shepard scatter (it moves only the last bits of the sums).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/sphdem_c.h"

#define NDIM 3
#define KERNAL_RADIUS 2.0 /* sphConstants.h:4-8, WENDLAND */
#define IAM_SPH 0
#define IAM_SPHB 1
#define IAM_BOUNDARY 3
#define IAM_GHOST 4
#define IAM_BOBJ1 5
#define IAM_BOBJ6 10
#define IAM_DEM 11
#define IAM_DEMB 12
#define IAM_IMMERSED 13

/* ---- particle record: the Cparticle fields the loop touches (particle.h:62-316) */
typedef struct Particle {
  double r[3], v[3], v0[3], vhat[3], vhat0[3];
  double f[3], fp[3], fv[3], fb[3], ff[3], fdrag[3], fvel[3];
  double mass, h, dens, dddt, press, pdr2, omega, porosity, shepSum, u, dudt;
  /* per-particle energies of calcEnergies (particle.h:62-316) */
  double eViscF, deViscFdt, eViscB, eBForce, eFF, deFFdt;
  int iam, tag;
  int shift[3]; /* ghosts: periodic image shift (-1,0,1) per axis; 0 for real */
  /* cached neighbour list (CpInfo::neighbrs, dataLL.h:76-84) */
  struct Particle** nb;
  int nnb, capnb;
} Particle;

typedef struct {
  Particle** e;
  int n, cap;
} PList;

typedef struct Oracle {
  sphdem_params P;
  /* ps: the particleContainer; gh: CdataLL::procGhostParticles */
  Particle* ps;
  int n, cap;
  Particle* gh;
  int ng, gcap;
  Particle** ghost_origin; /* ghostedParticles, in ghost creation order */
  int* ghost_dir;          /* direction (0..26) each ghost was sent to */
  /* CglobalVars */
  double time, dt, newDt, maxdt;
  long sphStep;
  double custom[9];
  /* CdataLL */
  double hmax, ssr;
  double procGhostMax2H[27];
  int cdim[3];
  PList* cells;
  int ddim[3];
  PList* dem_cells;
  long ndeleted;
} Oracle;

/* ------------------------------------------------------------------ helpers */
static double dot3(const double* a, const double* b) {
  /* blitz _bz_meta_vectorDot (blitz/meta/dot.h:40-60): a0*b0 + (a1*b1 + a2*b2) */
  return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}
static double len2v(const double* a) { return dot3(a, a); } /* vect.cpp:3-5 */
static double lenv(const double* a) { return sqrt(len2v(a)); } /* vect.cpp:7-9 */
static double dmin2(double a, double b) { return (b < a) ? b : a; } /* std::min */
static double dmax2(double a, double b) { return (a < b) ? b : a; } /* std::max */

static void plist_push(PList* l, Particle* p) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 8;
    l->e = (Particle**)realloc(l->e, sizeof(Particle*) * l->cap);
  }
  l->e[l->n++] = p;
}
static void nb_push(Particle* a, Particle* b) {
  if (a->nnb == a->capnb) {
    a->capnb = a->capnb ? 2 * a->capnb : 64;
    a->nb = (Particle**)realloc(a->nb, sizeof(Particle*) * a->capnb);
  }
  a->nb[a->nnb++] = b;
}

/* ------------------------------------------------- kernel family, WENDLAND */
/* F(q,h) sph.h:74-81 */
static double kF(const Oracle* o, double q, double h) {
  if (q <= 2.0)
    return (1 / pow(h, NDIM + 2)) * o->P.wcon *
           (-4 * pow(2 - q, 3) * (1 + 2 * q) + 2 * pow(2 - q, 4)) / q;
  return 0.0;
}
/* K(q,h) sph.h:199-205 */
static double kK(const Oracle* o, double q, double h) {
  (void)h;
  if (q <= 2.0) return o->P.wcon * pow(2.0 - q, 4) * (1.0 + 2.0 * q);
  return 0.0;
}
/* pdKh(q,h) sph.h:285-292 */
static double kpdKh(const Oracle* o, double q, double h) {
  if (q <= 2.0)
    return -o->P.wcon * q * (-4 * pow(2 - q, 3) * (1 + 2 * q) + 2 * pow(2 - q, 4)) / h;
  return 0.0;
}
/* W(q,h) sph.h:306-308 */
static double kW(const Oracle* o, double q, double h) { return (1 / pow(h, NDIM)) * kK(o, q, h); }
/* pdWh(q,h) sph.h:310-312 */
static double kpdWh(const Oracle* o, double q, double h) {
  return (1 / pow(h, NDIM)) * kpdKh(o, q, h) - NDIM / pow(h, NDIM + 1) * kK(o, q, h);
}
/* gradW(pa,pb,dr,q,h) sph.h:322-342 (no CORRECTED_GRADIENT) */
static void gradW(const Oracle* o, const Particle* pa, const double* dr, double q, double h,
                  double* out) {
  double drF[3] = {0.0, 0.0, 0.0};
  if (q != 0) {
    double F = kF(o, q, h);
    for (int i = 0; i < 3; i++) drF[i] = dr[i] * F;
  }
  if (o->P.flags & SPHDEM_F_VAR_H_CORRECTION) {
    for (int i = 0; i < 3; i++) out[i] = drF[i] / pa->omega;
  } else {
    for (int i = 0; i < 3; i++) out[i] = drF[i];
  }
}

/* ------------------------------------------------------------- predicates */
static int ifBoundary(const Particle* p) { /* sph.h:1882-1884 */
  return p->iam == IAM_BOUNDARY || (p->iam >= IAM_BOBJ1 && p->iam <= IAM_BOBJ6);
}
static int ifSphOrSphBoundary(const Particle* p) { /* sph.h:1860-1862 (and ...OrImmersedDem) */
  return p->iam == IAM_SPH || p->iam == IAM_SPHB;
}
static int ifDem(const Particle* p) { return p->iam == IAM_DEM; }

/* ------------------------------------------------------------ cell indices */
/* CdataLL::getCellI dataLL.h:303-305 */
static void getCellI(const Oracle* o, const double* r, int* c) {
  double e = KERNAL_RADIUS * o->ssr;
  for (int i = 0; i < 3; i++) c[i] = (int)((r[i] - o->P.proc_domain[2 * i]) / e + 3);
}
/* CdataLL::dem_getCellI dataLL.h:308-311 */
static void demGetCellI(const Oracle* o, const double* r, int* c) {
  double e = 2 * o->P.dem_search_radius;
  double s = 2 * o->P.coupling_radius / o->P.dem_search_radius;
  for (int i = 0; i < 3; i++) c[i] = (int)((r[i] - o->P.proc_domain[2 * i]) / e + s - 1);
}
static PList* cellAt(Oracle* o, const int* c) {
  for (int i = 0; i < 3; i++)
    if (c[i] < 0 || c[i] >= o->cdim[i]) return NULL;
  return &o->cells[((size_t)c[0] * o->cdim[1] + c[1]) * o->cdim[2] + c[2]];
}
static PList* demCellAt(Oracle* o, const int* c) {
  for (int i = 0; i < 3; i++)
    if (c[i] < 0 || c[i] >= o->ddim[i]) return NULL;
  return &o->dem_cells[((size_t)c[0] * o->ddim[1] + c[1]) * o->ddim[2] + c[2]];
}

/* Cparticle::is_neighbr particle.h:236-238 */
static int is_neighbr(const Particle* a, const Particle* b) {
  double d[3];
  for (int i = 0; i < 3; i++) d[i] = a->r[i] - b->r[i];
  return (a != b) && (len2v(d) <= pow(KERNAL_RADIUS * dmax2(a->h, b->h), 2));
}

/* cells.resize(...) dataLL.cpp:16, 222; dem_cells.resize dataLL.cpp:18 */
static void resizeCells(Oracle* o) {
  size_t nc = (size_t)o->cdim[0] * o->cdim[1] * o->cdim[2];
  if (o->cells) {
    for (size_t i = 0; i < nc; i++) free(o->cells[i].e);
    free(o->cells);
  }
  for (int i = 0; i < 3; i++)
    o->cdim[i] = (int)((o->P.proc_domain[2 * i + 1] - o->P.proc_domain[2 * i]) /
                       (KERNAL_RADIUS * o->ssr)) + 8;
  nc = (size_t)o->cdim[0] * o->cdim[1] * o->cdim[2];
  o->cells = (PList*)calloc(nc, sizeof(PList));
}
static void resizeDemCells(Oracle* o) {
  /* dataLL.cpp:18 sizes (cell_max-cell_min)/(2 DSR) + 4 Rc/DSR per axis; one
   * more cell on each side keeps every stencil access in range. */
  for (int i = 0; i < 3; i++)
    o->ddim[i] = (int)((o->P.proc_domain[2 * i + 1] - o->P.proc_domain[2 * i]) /
                       (2 * o->P.dem_search_radius)) +
                 (int)(2 * 2 * o->P.coupling_radius / o->P.dem_search_radius) + 2;
  size_t nc = (size_t)o->ddim[0] * o->ddim[1] * o->ddim[2];
  o->dem_cells = (PList*)calloc(nc, sizeof(PList));
}
static void clearCells(Oracle* o) {
  size_t nc = (size_t)o->cdim[0] * o->cdim[1] * o->cdim[2];
  for (size_t i = 0; i < nc; i++) o->cells[i].n = 0;
  if (o->dem_cells) {
    nc = (size_t)o->ddim[0] * o->ddim[1] * o->ddim[2];
    for (size_t i = 0; i < nc; i++) o->dem_cells[i].n = 0;
  }
}

/* CdataLL::insert dataLL.cpp:978-997 */
static int insertP(Oracle* o, Particle* p) {
  int c[3];
  PList* l;
  if ((o->P.flags & SPHDEM_F_LIQ_DEM) && p->iam == IAM_DEM) {
    demGetCellI(o, p->r, c);
    l = demCellAt(o, c);
  } else {
    getCellI(o, p->r, c);
    l = cellAt(o, c);
  }
  if (!l) {
    fprintf(stderr, "oracle: particle tag %d outside the cell grid\n", p->tag);
    return -1;
  }
  plist_push(l, p);
  return 0;
}

/* ----------------------------------------------------- domain / ghosts */
static int dirIndex(const int* c) { return (c[0] * 3 + c[1]) * 3 + c[2]; } /* blitz row-major */
static int hasNeighbr(const Oracle* o, const int* c) {
  /* globals.procNeighbrs of a 1x1x1 split (misc.cpp:690-720, setup.cpp:112-113):
   * >= 0 only through periodic faces, never for the centre. */
  if (c[0] == 1 && c[1] == 1 && c[2] == 1) return 0;
  for (int i = 0; i < 3; i++)
    if (c[i] != 1 && !o->P.periodic[i]) return 0;
  return 1;
}
static void applyPeriodicShift(const Oracle* o, const int* coords, double* r, int* shift) {
  /* dataLL.cpp:469-484 / 651-666 (DENS_DROP = 0, GHOST = 0) */
  for (int i = 0; i < 3; i++) {
    if (!o->P.periodic[i]) continue;
    if (coords[i] == 0) {
      r[i] = o->P.rmin[i] - o->P.rmax[i] + r[i];
      if (shift) shift[i] = -1;
    }
    if (coords[i] == 2) {
      r[i] = o->P.rmax[i] - o->P.rmin[i] + r[i];
      if (shift) shift[i] = 1;
    }
  }
}

static void clearNb(Particle* p) { p->nnb = 0; }

/* CdataLL::deleteParticle dataLL.cpp:285-313: the last particle fills the hole */
static void deleteParticle(Oracle* o, int k) {
  if (k != o->n - 1) {
    Particle** keep = o->ps[k].nb;
    int kc = o->ps[k].capnb;
    o->ps[k] = o->ps[o->n - 1];
    /* swap the list buffers so that each slot still owns one buffer */
    o->ps[o->n - 1].nb = keep;
    o->ps[o->n - 1].capnb = kc;
    o->ps[k].nnb = 0;
  }
  o->n--;
}

static void ghostCopy(Particle* g, const Particle* p) {
  /* Cparticle() defaults (particle.h:70-135) then operator=(CghostData)
   * (particle.h:322-398): r v vhat porosity shepSum fdrag mass dens h tag iam */
  memset(g, 0, sizeof(*g));
  g->porosity = 1.0;
  for (int i = 0; i < 3; i++) {
    g->r[i] = p->r[i];
    g->v[i] = p->v[i];
    g->vhat[i] = p->vhat[i];
    g->fdrag[i] = p->fdrag[i];
  }
  g->porosity = p->porosity;
  g->shepSum = p->shepSum;
  g->mass = p->mass;
  g->dens = p->dens;
  g->h = p->h;
  g->tag = p->tag;
  g->iam = p->iam;
}

/* CdataLL::updateDomain dataLL.cpp:319-700 for one rank whose only neighbours
 * are itself through periodic faces. */
static void updateDomain(Oracle* o) {
  double dmin[3], dspace[3], liq_dgmin[3], liq_dgspace[3], dem_dgmin[3], dem_dgspace[3];
  o->ng = 0;
  for (int i = 0; i < 3; i++) {
    int c[3] = {1, 1, 1};
    dmin[i] = o->P.proc_domain[2 * i];
    dspace[i] = o->P.proc_domain[2 * i + 1] - dmin[i];
    c[i] = 0;
    liq_dgmin[i] = dmin[i] + o->procGhostMax2H[dirIndex(c)];
    dem_dgmin[i] = dmin[i] + o->P.coupling_radius;
    liq_dgspace[i] = dspace[i] - o->procGhostMax2H[dirIndex(c)];
    dem_dgspace[i] = dspace[i] - o->P.coupling_radius;
    c[i] = 2;
    liq_dgspace[i] -= o->procGhostMax2H[dirIndex(c)];
    dem_dgspace[i] -= o->P.coupling_radius;
  }
  /* migrate: dataLL.cpp:370-403 */
  Particle* sendp[27];
  int nsendp[27], capsendp[27];
  for (int d = 0; d < 27; d++) {
    sendp[d] = NULL;
    nsendp[d] = capsendp[d] = 0;
  }
  for (int k = 0; k < o->n;) {
    Particle* p = &o->ps[k];
    clearNb(p);
    if (p->tag == -111) {
      deleteParticle(o, k);
      continue;
    }
    int oc[3], out = 0, notone = 0;
    for (int i = 0; i < 3; i++) {
      oc[i] = (int)((p->r[i] - dmin[i]) / dspace[i] + 1);
      if (oc[i] > 2 || oc[i] < 0) out = 1;
      if (oc[i] != 1) notone = 1;
    }
    if (out) {
      o->ndeleted++;
      deleteParticle(o, k);
      continue;
    }
    if (notone) {
      if (hasNeighbr(o, oc)) {
        int d = dirIndex(oc);
        if (nsendp[d] == capsendp[d]) {
          capsendp[d] = capsendp[d] ? 2 * capsendp[d] : 64;
          sendp[d] = (Particle*)realloc(sendp[d], sizeof(Particle) * capsendp[d]);
        }
        sendp[d][nsendp[d]] = *p;
        sendp[d][nsendp[d]].nb = NULL;
        sendp[d][nsendp[d]].nnb = sendp[d][nsendp[d]].capnb = 0;
        nsendp[d]++;
      } else {
        o->ndeleted++;
      }
      deleteParticle(o, k);
      continue;
    }
    k++;
  }
  /* receive, in blitz iteration order over procNeighbrs: dataLL.cpp:459-497.
   * Slot `coords` receives what was sent towards 2-coords. */
  for (int c0 = 0; c0 < 3; c0++)
    for (int c1 = 0; c1 < 3; c1++)
      for (int c2 = 0; c2 < 3; c2++) {
        int coords[3] = {c0, c1, c2};
        if (!hasNeighbr(o, coords)) continue;
        int opp[3] = {2 - c0, 2 - c1, 2 - c2};
        int d = dirIndex(opp);
        for (int z = 0; z < nsendp[d]; z++) {
          if (o->n >= o->cap) {
            fprintf(stderr, "oracle: particle capacity exceeded\n");
            abort();
          }
          Particle* np = &o->ps[o->n];
          Particle** keep = np->nb;
          int kc = np->capnb;
          *np = sendp[d][z];
          np->nb = keep;
          np->capnb = kc;
          np->nnb = 0;
          applyPeriodicShift(o, coords, np->r, NULL);
          o->n++;
        }
      }
  for (int d = 0; d < 27; d++) free(sendp[d]);

  /* ghosts to send: dataLL.cpp:503-583 */
  int* gsend[27];
  int ngsend[27], capgsend[27];
  for (int d = 0; d < 27; d++) {
    gsend[d] = NULL;
    ngsend[d] = capgsend[d] = 0;
  }
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    clearNb(p);
    int ic[3];
    for (int i = 0; i < 3; i++) {
      if ((o->P.flags & SPHDEM_F_LIQ_DEM) && p->iam == IAM_DEM)
        ic[i] = (int)((p->r[i] - dem_dgmin[i]) / dem_dgspace[i] + 1);
      else
        ic[i] = (int)((p->r[i] - liq_dgmin[i]) / liq_dgspace[i] + 1);
      if (ic[i] > 2) ic[i] = 2;
      if (ic[i] < 0) ic[i] = 0;
    }
    if (ic[0] == 1 && ic[1] == 1 && ic[2] == 1) continue;
    int lo[3], hi[3];
    for (int i = 0; i < 3; i++) {
      if (ic[i] <= 1) {
        hi[i] = 1;
        lo[i] = ic[i];
      } else {
        hi[i] = ic[i];
        lo[i] = 1;
      }
    }
    for (int a = lo[0]; a <= hi[0]; a++)
      for (int b = lo[1]; b <= hi[1]; b++)
        for (int c = lo[2]; c <= hi[2]; c++) {
          int dc[3] = {a, b, c};
          if (!hasNeighbr(o, dc)) continue;
          int d = dirIndex(dc);
          if (ngsend[d] == capgsend[d]) {
            capgsend[d] = capgsend[d] ? 2 * capgsend[d] : 64;
            gsend[d] = (int*)realloc(gsend[d], sizeof(int) * capgsend[d]);
          }
          gsend[d][ngsend[d]++] = k;
        }
  }
  /* receive ghosts: dataLL.cpp:631-695 */
  for (int c0 = 0; c0 < 3; c0++)
    for (int c1 = 0; c1 < 3; c1++)
      for (int c2 = 0; c2 < 3; c2++) {
        int coords[3] = {c0, c1, c2};
        if (!hasNeighbr(o, coords)) continue;
        int opp[3] = {2 - c0, 2 - c1, 2 - c2};
        int d = dirIndex(opp);
        double maxh = o->hmax;
        for (int z = 0; z < ngsend[d]; z++) {
          if (o->ng >= o->gcap) {
            o->gcap = o->gcap ? 2 * o->gcap : 1024;
            o->gh = (Particle*)realloc(o->gh, sizeof(Particle) * o->gcap);
            o->ghost_origin = (Particle**)realloc(o->ghost_origin, sizeof(Particle*) * o->gcap);
            o->ghost_dir = (int*)realloc(o->ghost_dir, sizeof(int) * o->gcap);
          }
          Particle* g = &o->gh[o->ng];
          ghostCopy(g, &o->ps[gsend[d][z]]);
          applyPeriodicShift(o, coords, g->r, g->shift);
          if (g->h > maxh) maxh = g->h;
          o->ghost_origin[o->ng] = &o->ps[gsend[d][z]];
          o->ghost_dir[o->ng] = d;
          o->ng++;
        }
        o->procGhostMax2H[dirIndex(coords)] = KERNAL_RADIUS * maxh;
      }
  for (int d = 0; d < 27; d++) free(gsend[d]);
}

/* CdataLL::calcDomainLimits dataLL.cpp:954-975 (Morton bounds not needed) */
static void calcDomainLimits(Oracle* o) {
  o->hmax = -100000;
  for (int k = 0; k < o->n; k++)
    if (o->ps[k].h > o->hmax) o->hmax = o->ps[k].h;
  for (int i = 0; i < 3; i++) {
    int c[3] = {1, 1, 1};
    c[i] = 0;
    if (o->procGhostMax2H[dirIndex(c)] < o->hmax) o->procGhostMax2H[dirIndex(c)] = KERNAL_RADIUS * o->hmax;
    c[i] = 2;
    if (o->procGhostMax2H[dirIndex(c)] < o->hmax) o->procGhostMax2H[dirIndex(c)] = KERNAL_RADIUS * o->hmax;
  }
}

/* CdataLL::reset dataLL.cpp:140-236 */
static int resetO(Oracle* o) {
  clearCells(o);
  updateDomain(o);
  calcDomainLimits(o);
  if ((o->hmax > o->ssr) || (o->hmax < 0.9 * o->ssr)) {
    o->ssr = o->hmax;
    resizeCells(o);
  }
  for (int k = 0; k < o->n; k++)
    if (insertP(o, &o->ps[k])) return -1;
  for (int k = 0; k < o->ng; k++)
    if (insertP(o, &o->gh[k])) return -1;
  return 0;
}

/* ------------------------------------------------------ neighbour searches */
static void addIfNeighbr(Particle* p, PList* l, Particle** out_owner) {
  (void)out_owner;
  if (!l) return;
  for (int i = 0; i < l->n; i++)
    if (is_neighbr(p, l->e[i])) nb_push(p, l->e[i]);
}

/* CdataLL::calcNeighbours dataLL.cpp:1026-1075: the list is stored in p->nb */
static void calcNeighbours(Oracle* o, Particle* p) {
  int c[3], t[3];
  p->nnb = 0;
  if ((o->P.flags & SPHDEM_F_LIQ_DEM) && p->iam == IAM_DEM) {
    demGetCellI(o, p->r, c);
    for (t[0] = c[0] - 1; t[0] <= c[0] + 1; t[0]++)
      for (t[1] = c[1] - 1; t[1] <= c[1] + 1; t[1]++)
        for (t[2] = c[2] - 1; t[2] <= c[2] + 1; t[2]++) addIfNeighbr(p, demCellAt(o, t), NULL);
  } else {
    getCellI(o, p->r, c);
    for (t[0] = c[0] - 1; t[0] <= c[0] + 1; t[0]++)
      for (t[1] = c[1] - 1; t[1] <= c[1] + 1; t[1]++)
        for (t[2] = c[2] - 1; t[2] <= c[2] + 1; t[2]++) addIfNeighbr(p, cellAt(o, t), NULL);
  }
}
/* pInfo neighbour cache: lists are (re)built when empty (dataLL.impl.h:186-188) */
static void ensureNeighbours(Oracle* o, Particle* p) {
  if (p->nnb == 0) calcNeighbours(o, p);
}

/* CdataLL::calcNeighboursCoupling dataLL.cpp:1104-1192 into l (not cached) */
static void calcNeighboursCoupling(Oracle* o, Particle* p, PList* out) {
  out->n = 0;
  if (p->iam == IAM_DEM) {
    int c[3], t[3];
    getCellI(o, p->r, c);
    for (t[0] = c[0] - 1; t[0] <= c[0] + 1; t[0]++)
      for (t[1] = c[1] - 1; t[1] <= c[1] + 1; t[1]++)
        for (t[2] = c[2] - 1; t[2] <= c[2] + 1; t[2]++) {
          PList* l = cellAt(o, t);
          if (!l) continue;
          for (int i = 0; i < l->n; i++)
            if (is_neighbr(p, l->e[i])) plist_push(out, l->e[i]);
        }
  } else {
    const double Rc = o->P.coupling_radius, DSR = o->P.dem_search_radius;
    const int num_cells = (int)ceil(0.5 * Rc / DSR);
    const double thr = pow(0.5 * Rc / DSR, 2);
    const double rc2 = pow(Rc, 2);
    int c[3], t[3];
    demGetCellI(o, p->r, c);
    for (t[0] = c[0] - num_cells; t[0] <= c[0] + num_cells; t[0]++) {
      const int iPlus2 = (int)pow(t[0] - c[0] + 1, 2);
      const int iMinus2 = (int)pow(t[0] - c[0] - 1, 2);
      for (t[1] = c[1] - num_cells; t[1] <= c[1] + num_cells; t[1]++) {
        const int jPlus2 = (int)pow(t[1] + 1 - c[1], 2);
        const int jMinus2 = (int)pow(t[1] - 1 - c[1], 2);
        for (t[2] = c[2] - num_cells; t[2] <= c[2] + num_cells; t[2]++) {
          const int rPlus2 = (int)(iPlus2 + jPlus2 + pow(t[2] - c[2] + 1, 2));
          const int rMinus2 = (int)(iMinus2 + jMinus2 + pow(t[2] - c[2] - 1, 2));
          PList* l = demCellAt(o, t);
          if (!l) continue;
          if (rPlus2 < thr) {
            for (int i = 0; i < l->n; i++) /* addAll dataLL.cpp:1018-1023 */
              if (p != l->e[i]) plist_push(out, l->e[i]);
          } else if (rMinus2 < thr) {
            for (int i = 0; i < l->n; i++) { /* addIfInRadius dataLL.cpp:1009-1016 */
              double d[3];
              for (int k = 0; k < 3; k++) d[k] = p->r[k] - l->e[i]->r[k];
              if ((len2v(d) < rc2) && (p != l->e[i])) plist_push(out, l->e[i]);
            }
          }
        }
      }
    }
  }
}

/* -------------------------------------------------------- pair operators */
/* calcViscForce sph.h:579-666 */
static void calcViscForce(Oracle* o, Particle* pa, Particle* pb, const double* gradWa,
                          const double* dv, double r, double* newDt) {
  const sphdem_params* P = &o->P;
  double dx[3], fv[3] = {0, 0, 0};
  for (int i = 0; i < 3; i++) dx[i] = pa->r[i] - pb->r[i];
  double vdr = dot3(dx, dv);
  double hav = 0.5 * (pa->h + pb->h);
  double vsig = 0.0;
  if (P->flags & SPHDEM_F_VISC_MORRIS) {
    double visc = P->viscosity * (pa->dens + pb->dens) / (pa->dens * pb->dens);
    vsig = 2.0 * P->spsound;
    if (P->flags & SPHDEM_F_LIQ_DEM) visc = 2.0 * P->viscosity * P->dens / (pa->dens * pb->dens);
    double d = dot3(dx, gradWa);
    double den = pow(r, 2) + 0.01 * pow(hav, 2);
    for (int i = 0; i < 3; i++) fv[i] = dv[i] * pb->mass * visc * d / den;
  }
  if (P->flags & SPHDEM_F_VISC_CLEARY) { /* sph.h:598-604 */
    double visc = 19.8 * P->viscosity * vdr / ((pa->dens + pb->dens) * (pow(r, 2) + 0.01 * pow(hav, 2)));
    vsig = 2.0 * P->spsound;
    for (int i = 0; i < 3; i++) fv[i] = pb->mass * visc * gradWa[i];
  }
  if (P->flags & SPHDEM_F_VISC_MONAGHAN) { /* sph.h:606-630 (no SLK) */
    double dr = 0.0;
    if (r != 0.0) dr = 1 / r;
    double viss = vdr * dr;
    /* under LIQ_DEM the signal speed stays 0 (its replacement is commented out);
     * abs(viss) is the double overload (gsl_math.h includes <math.h>, whose
     * C++ form declares std::abs in the global namespace) */
    if (!(P->flags & SPHDEM_F_LIQ_DEM)) vsig = 2.0 * P->spsound + 2.0 * fabs(viss);
    /* ALPHA = viscToAlpha(VISCOSITY, H, SPSOUND) (sph.h:42, misc.h:41-43) */
    const double alpha = P->viscosity * (112.0 / 15.0) / (P->spsound * P->h);
    double visc = viss * vsig * alpha / (pa->dens + pb->dens);
    for (int i = 0; i < 3; i++) fv[i] = pb->mass * visc * gradWa[i];
  }
  if (P->flags & SPHDEM_F_VISC_ARTIFICIAL) {
    double artificial_vsig = 0, artificial_visc = 0.0;
    if (vdr < 0) {
      double dr = 0.0;
      if (r != 0.0) dr = 1 / r;
      double viss = vdr * dr;
      artificial_vsig = 2.0 * P->spsound - 4.0 * viss;
      artificial_visc = viss * artificial_vsig * P->alpha_artificial / (pa->dens + pb->dens);
    }
    double s = pb->mass * artificial_visc;
    for (int i = 0; i < 3; i++) pa->f[i] += s * gradWa[i];
    vsig = dmax2(vsig, artificial_vsig);
  }
  *newDt = dmin2(*newDt, 0.6 * hav / vsig);                            /* sph.h:545-552 */
  /* viscDiffusionCondition sph.h:554-560 */
  const double cdiff = (P->flags & SPHDEM_F_VISC_MONAGHAN) ? 0.06 : 0.125;
  *newDt = dmin2(*newDt, cdiff * pow(hav, 2) / P->viscosity);
  for (int i = 0; i < 3; i++) pa->fv[i] += fv[i];
  for (int i = 0; i < 3; i++) pa->f[i] += fv[i];
}

/* calcPressForce sph.h:692-734 (kdwPowNeps = 0 under NO_ANTICLUMPING) */
static void calcPressForce(Particle* pa, Particle* pb, const double* gWa, const double* gWb) {
  const double kdwPowNeps = 0.0;
  double antic = 0.01 * fabs(pa->pdr2 + pb->pdr2) * kdwPowNeps;
  for (int i = 0; i < 3; i++) {
    double prfac = pa->pdr2 * gWa[i] + pb->pdr2 * gWb[i];
    prfac = prfac + antic;
    double fp = -pb->mass * prfac;
    pa->fp[i] += fp;
    pa->f[i] += fp;
  }
}

/* CsphIncompress::calcForce sphIncompress.h:899-972 */
static void calcForce(Oracle* o, Particle* pa, Particle* pb, double* newDt) {
  double dx[3], gWa[3], gWb[3];
  for (int i = 0; i < 3; i++) dx[i] = pa->r[i] - pb->r[i];
  const double r = lenv(dx);
  if (o->P.flags & SPHDEM_F_VAR_H_CORRECTION) {
    gradW(o, pa, dx, r / pa->h, pa->h, gWa);
    gradW(o, pa, dx, r / pb->h, pb->h, gWb);
  } else {
    double g1[3], g2[3];
    gradW(o, pa, dx, r / pa->h, pa->h, g1);
    gradW(o, pa, dx, r / pb->h, pb->h, g2);
    for (int i = 0; i < 3; i++) gWa[i] = gWb[i] = 0.5 * (g1[i] + g2[i]);
  }
  if (pa->iam == IAM_SPH && pb->iam != IAM_DEM && pb->iam != IAM_DEMB) {
    double dv[3], gbar[3];
    for (int i = 0; i < 3; i++) dv[i] = pa->vhat[i] - pb->vhat[i];
    for (int i = 0; i < 3; i++) gbar[i] = 0.5 * (gWa[i] + gWb[i]);
    calcViscForce(o, pa, pb, gbar, dv, r, newDt);
  }
  if (ifBoundary(pb)) {
    /* calcBoundaryForces (type boundary walls) is outside the supported flag set */
  } else if (pb->iam == IAM_SPH || pb->iam == IAM_SPHB) {
    calcPressForce(pa, pb, gWa, gWb);
  }
}

/* CsphIncompress::calcDddtDudt sphIncompress.h:783-851 */
static void calcDddtDudt(Oracle* o, Particle* pa, Particle* pb) {
  if (ifBoundary(pb) || pb->iam == IAM_DEM) return;
  double dr[3], dv[3], gWa[3];
  for (int i = 0; i < 3; i++) {
    dr[i] = pa->r[i] - pb->r[i];
    dv[i] = pa->vhat[i] - pb->vhat[i];
  }
  const double r = lenv(dr);
  if (o->P.flags & SPHDEM_F_VAR_H_CORRECTION) {
    gradW(o, pa, dr, r / pa->h, pa->h, gWa);
  } else {
    double g1[3], g2[3];
    gradW(o, pa, dr, r / pa->h, pa->h, g1);
    gradW(o, pa, dr, r / pb->h, pb->h, g2);
    for (int i = 0; i < 3; i++) gWa[i] = 0.5 * (g1[i] + g2[i]);
  }
  const double vdr = dot3(dv, gWa);
  const double dddtInc = pb->mass * vdr;
  pa->dddt += dddtInc;
}

/* Nsph::calcOmega sph.h:437-452 */
static void calcOmega(Oracle* o, Particle* p) {
  p->omega = p->mass * kpdWh(o, 0, p->h);
  for (int i = 0; i < p->nnb; i++) {
    Particle* pn = p->nb[i];
    if (pn->iam == IAM_SPH || pn->iam == IAM_SPHB) {
      double dr[3];
      for (int k = 0; k < 3; k++) dr[k] = p->r[k] - pn->r[k];
      double r = lenv(dr);
      p->omega += pn->mass * kpdWh(o, r / p->h, p->h);
    }
  }
  p->omega = 1 + p->h * p->omega / (p->dens * NDIM);
}

/* Nsph::calcDensity sph.h:1149-1191 (no REINIT_DENS_MLS) */
static void calcDensity(Oracle* o, Particle* p) {
  p->dens = p->mass * kW(o, 0, p->h);
  for (int i = 0; i < p->nnb; i++) {
    Particle* pn = p->nb[i];
    if ((pn->iam != IAM_SPH) && (pn->iam != IAM_SPHB)) continue;
    double dr[3];
    for (int k = 0; k < 3; k++) dr[k] = p->r[k] - pn->r[k];
    double r = lenv(dr);
    double Wa = kW(o, r / p->h, p->h);
    p->dens += pn->mass * Wa;
  }
}

/* Nsph::calcHFromDens sph.h:1823-1825 */
static void calcHFromDens(Oracle* o, Particle* p) { p->h = o->P.hfac * pow(p->mass / p->dens, 1.0 / NDIM); }

/* CsphIncompress::calcPressSpsoundPdr2 sphIncompress.h:863-896 */
static void calcPressSpsoundPdr2(Oracle* o, Particle* p) {
  const sphdem_params* P = &o->P;
  if (P->flags & SPHDEM_F_LIQ_DEM)
    p->press = P->prb * (pow(p->dens / (P->refd * p->porosity), P->gamma) - 1.0);
  else
    p->press = P->prb * (pow(p->dens / P->refd, 7) - 1.0);
  p->pdr2 = p->press / (p->dens * p->dens);
}

/* ------------------------------------------------------- coupling operators */
/* calcPorosity sph.h:1323-1341 with calcPorosityIncr sph.h:1246-1256 */
static void calcPorosity(Oracle* o, Particle* p, PList* nb) {
  const double Rc = o->P.coupling_radius;
  p->porosity = 0;
  for (int i = 0; i < nb->n; i++) {
    Particle* pn = nb->e[i];
    double dx[3];
    for (int k = 0; k < 3; k++) dx[k] = p->r[k] - pn->r[k];
    const double r = lenv(dx);
    const double q = r * (2.0 / Rc);
    const double Wa = kW(o, q, Rc / 2.0);
    p->porosity += o->P.porosity_incr * Wa;
  }
}
/* calcShepSum sph.h:1303-1320: scatter into the DEM partners */
static void calcShepSum(Oracle* o, Particle* p, PList* nb) {
  const double Rc = o->P.coupling_radius;
  double dv = p->mass / p->dens;
  for (int i = 0; i < nb->n; i++) {
    Particle* pn = nb->e[i];
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = p->r[k] - pn->r[k];
    const double r = lenv(d);
    pn->shepSum += dv * kW(o, r * (2.0 / Rc), Rc / 2.0);
  }
}
/* calcDEMtoSPHDrag sph.h:1343-1362 */
static void calcDEMtoSPHDrag(Oracle* o, Particle* p, PList* nb) {
  const double Rc = o->P.coupling_radius;
  for (int k = 0; k < 3; k++) p->fdrag[k] = 0.0;
  const double rdens = 1.0 / p->dens;
  for (int i = 0; i < nb->n; i++) {
    Particle* pn = nb->e[i];
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = p->r[k] - pn->r[k];
    const double r = lenv(d);
    const double s = pn->mass * rdens;
    const double w = kW(o, r * (2.0 / Rc), Rc / 2.0);
    for (int k = 0; k < 3; k++) p->fdrag[k] -= s * pn->fdrag[k] * w / pn->shepSum;
  }
  for (int k = 0; k < 3; k++) p->f[k] += p->fdrag[k];
}
/* finalisePorosity sph.h:1289-1296 */
static void finalisePorosity(Particle* p) {
  p->porosity = 1.0 - p->porosity;
  if (p->porosity < 0.37) p->porosity = 0.37;
}

/* calcParticleDrag sph.h:1575-1631 (Di Felice branch) */
static void calcParticleDrag(Oracle* o, Particle* p, const double* fvel, double* out) {
  const sphdem_params* P = &o->P;
  double dv[3], fd[3] = {0, 0, 0};
  for (int i = 0; i < 3; i++) dv[i] = p->porosity * (fvel[i] - p->vhat[i]);
  const double vdiff = lenv(dv);
  double pmass = p->mass;
  if (vdiff != 0) {
    const double A = P->pi * pow(P->dem_radius, 2);
    const double Re = (2.0 * P->dem_radius / P->viscosity) * vdiff;
    const double dragGamma = P->drag_a - P->drag_b * exp(-pow(1.5 - log10(Re), 2) / 2.0);
    const double C = pow(0.63 + 4.8 / pow(Re, 0.5), 2);
    const double s = (1.0 / 2.0) * A * C * vdiff;
    const double pw = pow(p->porosity, -dragGamma);
    for (int i = 0; i < 3; i++) fd[i] = s * dv[i] * pw;
  }
  for (int i = 0; i < 3; i++) out[i] = (P->dens / pmass) * fd[i];
}

/* calcFPI sph.h:1708-1792 (non-TEST branch; fvort feeds only LIQ_DEM_LIFT) */
static void calcFPI(Oracle* o, Particle* p, PList* nb) {
  double sum = 0, fdens = 0;
  double ff[3] = {0, 0, 0}, fpfb[3] = {0, 0, 0}, fv[3] = {0, 0, 0}, fdrag[3] = {0, 0, 0};
  for (int k = 0; k < 3; k++) {
    p->fvel[k] = 0.0;
    p->fdrag[k] = 0.0;
  }
  p->porosity = 0.0;
  p->shepSum = 0.0;
  for (int i = 0; i < nb->n; i++) {
    Particle* pn = nb->e[i];
    if (pn->iam == IAM_BOUNDARY) continue;
    double dx[3];
    for (int k = 0; k < 3; k++) dx[k] = p->r[k] - pn->r[k];
    const double r = lenv(dx);
    const double dvol = pn->mass / pn->dens;
    const double q = r / pn->h;
    const double Wab = kW(o, q, pn->h);
    const double dvWab = dvol * Wab;
    const double mW = pn->mass * Wab;
    for (int k = 0; k < 3; k++) {
      fpfb[k] += mW * (pn->fp[k] + pn->fb[k]);
      fv[k] += mW * (pn->fv[k]);
    }
    for (int k = 0; k < 3; k++) {
      fv[k] += dvWab * (pn->fv[k]);
      ff[k] += dvWab * (pn->f[k]);
      fdrag[k] += dvWab * pn->fdrag[k];
      p->fvel[k] += dvWab * pn->vhat[k];
    }
    p->porosity += dvWab * pn->porosity;
    fdens += Wab * pn->mass;
    sum += dvWab;
  }
  p->shepSum = sum;
  if (p->shepSum > 0.5) {
    const double rsum = 1.0 / sum;
    for (int k = 0; k < 3; k++) {
      ff[k] *= rsum;
      fpfb[k] *= rsum;
      fv[k] *= rsum;
      fdrag[k] *= rsum;
      p->fvel[k] *= rsum;
    }
    fdens *= rsum;
    p->porosity *= rsum;
    for (int k = 0; k < 3; k++) {
      p->fb[k] = fpfb[k];
      p->fv[k] = fv[k];
    }
    double dvFdrag[3];
    calcParticleDrag(o, p, p->fvel, dvFdrag);
    const double s = o->P.dem_vol / p->mass;
    for (int k = 0; k < 3; k++) p->fdrag[k] = s * (fv[k] + fpfb[k]) + dvFdrag[k];
    for (int k = 0; k < 3; k++) p->f[k] += p->fdrag[k];
  }
}

/* demNormal (LINEAR) sph.h:1495-1516 */
static double demNormalScale(const Oracle* o, double overlap, double overlap_dot) {
  return o->P.dem_k * overlap + o->P.dem_gamma * overlap_dot;
}
/* calcDEMContacts sph.h:1676-1706 */
static void calcDEMContacts(Oracle* o, Particle* p) {
  for (int i = 0; i < p->nnb; i++) {
    Particle* pn = p->nb[i];
    double dx[3], dv[3], normal[3];
    for (int k = 0; k < 3; k++) {
      dx[k] = p->r[k] - pn->r[k];
      dv[k] = p->vhat[k] - pn->vhat[k];
    }
    const double r = lenv(dx);
    for (int k = 0; k < 3; k++) normal[k] = dx[k] / r;
    const double overlap = 2.0 * o->P.dem_radius - r;
    if (overlap > 0) {
      const double overlap_dot = -dot3(dv, normal);
      const double pmass = p->mass;
      const double s = demNormalScale(o, overlap, overlap_dot);
      for (int k = 0; k < 3; k++) {
        const double fN = s * normal[k] / pmass;
        p->fp[k] += fN;
        p->f[k] += fN;
      }
    }
  }
}

/* ----------------------------------------------------- integrator operators */
/* driftR sphIncompress.h:620-641 */
static void driftR(Oracle* o, Particle* p) {
  const double dt = o->dt / 2;
  for (int i = 0; i < 3; i++) p->r[i] += dt * p->vhat[i];
}
/* driftRAndKick sphIncompress.h:643-676 */
static void driftRAndKick(Oracle* o, Particle* p) {
  const double dt = o->dt / 2;
  for (int i = 0; i < 3; i++) p->r[i] += dt * p->vhat[i];
  if (p->iam != IAM_SPHB) {
    for (int i = 0; i < 3; i++) {
      p->v0[i] = p->v[i];
      p->vhat0[i] = p->vhat[i];
      p->v[i] += dt * p->f[i];
      p->vhat[i] += dt * p->f[i];
    }
    if ((o->P.flags & SPHDEM_F_FIXED_DEM) && p->iam == IAM_DEM) {
      for (int i = 0; i < 3; i++) {
        p->v[i] = p->v0[i];
        p->vhat[i] = p->vhat0[i];
      }
    }
  }
}
/* driftRest sphIncompress.h:678-704 */
static void driftRest(Oracle* o, Particle* p) {
  const double dt = o->dt / 2;
  p->dens += dt * p->dddt;
  if (!(o->P.flags & SPHDEM_F_CONST_H)) p->h = o->P.hfac * pow(p->mass / p->dens, 1.0 / NDIM);
}
/* kick sphIncompress.h:706-742 */
static void kick(Oracle* o, Particle* p, double halfDt0, double halfDt1) {
  double dt = halfDt0 + halfDt1;
  if ((p->iam == IAM_SPH) || (p->iam == IAM_DEM) || (p->iam == IAM_IMMERSED)) {
    for (int i = 0; i < 3; i++) {
      p->v[i] = p->v0[i] + dt * p->f[i];
      p->vhat[i] = p->vhat0[i] + dt * p->f[i];
    }
    if (o->time < o->P.damptime) {
      for (int i = 0; i < 3; i++) {
        p->v[i] *= 0.98;
        p->vhat[i] *= 0.98;
      }
    }
    if ((o->P.flags & SPHDEM_F_FIXED_DEM) && p->iam == IAM_DEM) {
      for (int i = 0; i < 3; i++) {
        p->v[i] = p->v0[i];
        p->vhat[i] = p->vhat0[i];
      }
    }
  }
}
/* initSumsMiddle sphIncompress.h:744-763 */
static void initSumsMiddle(Particle* p) {
  for (int i = 0; i < 3; i++) p->f[i] = p->fp[i] = p->fv[i] = p->fb[i] = p->ff[i] = 0.0;
  p->dudt = 0.0;
}
/* initSumsEnd sphIncompress.h:765-767 */
static void initSumsEnd(Particle* p) { p->dddt = 0.0; }
/* calcEnergies sphIncompress.h:769-781 */
static void calcEnergies(Oracle* o, Particle* p) {
  if (p->iam == IAM_SPH) {
    p->eViscB = 0;
    p->eBForce = 0;
    p->deViscFdt = -dot3(p->vhat, p->fv);
    p->eViscF += o->dt * p->deViscFdt;
    p->deFFdt = -dot3(p->vhat, p->ff);
    p->eFF += o->dt * p->deFFdt;
  }
  calcPressSpsoundPdr2(o, p);
  p->dudt = p->pdr2 * p->dddt;
  p->u += o->dt * p->dudt;
}
/* Nsph::accelCondition sph.h:566-568 */
static void accelCondition(Particle* p, double* newDt) {
  *newDt = dmin2(*newDt, 0.25 * sqrt(p->h / lenv(p->f)));
}

/* ---------------------------------------------------------- pass drivers */
static int isLiqDem(const Oracle* o) { return (o->P.flags & SPHDEM_F_LIQ_DEM) != 0; }

/* Pass drivers.  The gathers run over the particles with OpenMP (when built
 * with it): each particle's sums keep the reference's order, so the results
 * are those of a serial sweep; the Shepard scatter stays serial and the dt
 * minimum is exact in any order. */
static void passOmega(Oracle* o) { /* neighboursGroup<calcOmega,ifSph..> */
#pragma omp parallel for schedule(dynamic, 256)
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    ensureNeighbours(o, p);
    calcOmega(o, p);
  }
}
static void syncOmega(Oracle* o) { /* syncParticlesBetweenProcs<readOmega,...> */
  for (int g = 0; g < o->ng; g++) o->gh[g].omega = o->ghost_origin[g]->omega;
}
static void passEos(Oracle* o) { /* traverse<calcPressSpsoundPdr2,...OrGhost> */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) calcPressSpsoundPdr2(o, &o->ps[k]);
  for (int g = 0; g < o->ng; g++)
    if (ifSphOrSphBoundary(&o->gh[g])) calcPressSpsoundPdr2(o, &o->gh[g]);
}
static void passForce(Oracle* o) { /* neighbours<calcForce,ifSphOrSphBoundary> */
  double newDt = o->newDt;
#pragma omp parallel for schedule(dynamic, 256) reduction(min : newDt)
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    ensureNeighbours(o, p);
    for (int i = 0; i < p->nnb; i++) calcForce(o, p, p->nb[i], &newDt);
  }
  o->newDt = newDt;
}
static void passDddt(Oracle* o) { /* neighbours<calcDddtDudt,...> */
#pragma omp parallel for schedule(dynamic, 256)
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    ensureNeighbours(o, p);
    for (int i = 0; i < p->nnb; i++) calcDddtDudt(o, p, p->nb[i]);
  }
}
static void passDensityOnce(Oracle* o) { /* one iteration of sphIncompress.cpp:29-34 */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) calcHFromDens(o, &o->ps[k]);
  for (int g = 0; g < o->ng; g++)
    if (ifSphOrSphBoundary(&o->gh[g])) calcHFromDens(o, &o->gh[g]);
  for (int g = 0; g < o->ng; g++) o->gh[g].h = o->ghost_origin[g]->h;
#pragma omp parallel for schedule(dynamic, 256)
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    ensureNeighbours(o, p);
    calcDensity(o, p);
  }
  for (int g = 0; g < o->ng; g++) o->gh[g].dens = o->ghost_origin[g]->dens;
}
/* sphIncompress.cpp:64-78: porosity, Shepard sum, DEM->SPH drag, finalise */
static void passPorosity(Oracle* o) {
  /* The SPH->DEM search of each SPH particle (it depends on positions only,
   * which none of the three sweeps below changes), made once, in parallel;
   * each sweep then reads the particle's list in the search's order. */
  PList* cl = (PList*)calloc((size_t)o->n, sizeof(PList));
#pragma omp parallel for schedule(dynamic, 256)
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    calcNeighboursCoupling(o, p, &cl[k]);
    calcPorosity(o, p, &cl[k]);
  }
  /* traverse<initShepSum,ifDemOrGhost>: DEM and DEM ghosts */
  for (int k = 0; k < o->n; k++)
    if (ifDem(&o->ps[k])) o->ps[k].shepSum = 0.0;
  for (int g = 0; g < o->ng; g++)
    if (ifDem(&o->gh[g])) o->gh[g].shepSum = 0.0;
  /* the scatter, serial in container order */
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    calcShepSum(o, p, &cl[k]);
  }
  /* reverseSyncParticlesBetweenProcs<readShepSum,sumShepSum> dataLL.impl.h:414-477:
   * ghost values return to their origin, summed in direction order. */
  for (int d = 0; d < 27; d++)
    for (int g = 0; g < o->ng; g++)
      if (o->ghost_dir[g] == d) o->ghost_origin[g]->shepSum += o->gh[g].shepSum;
  for (int g = 0; g < o->ng; g++) o->gh[g].shepSum = o->ghost_origin[g]->shepSum;
#pragma omp parallel for schedule(dynamic, 256)
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    calcDEMtoSPHDrag(o, p, &cl[k]);
  }
  for (int k = 0; k < o->n; k++) free(cl[k].e);
  free(cl);
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) finalisePorosity(&o->ps[k]);
  for (int k = 0; k < o->n; k++)
    if (ifBoundary(&o->ps[k])) o->ps[k].porosity = 1.0; /* setPorosityForBoundary */
  for (int g = 0; g < o->ng; g++) o->gh[g].porosity = o->ghost_origin[g]->porosity;
}
static void passFpi(Oracle* o) { /* sync f fb fp fv, then neighboursGroupCoupling<calcFPI,ifDem> */
  for (int g = 0; g < o->ng; g++) {
    Particle* src = o->ghost_origin[g];
    for (int i = 0; i < 3; i++) {
      o->gh[g].f[i] = src->f[i];
      o->gh[g].fb[i] = src->fb[i];
      o->gh[g].fp[i] = src->fp[i];
      o->gh[g].fv[i] = src->fv[i];
    }
  }
#pragma omp parallel
  {
    PList nb = {0, 0, 0};
#pragma omp for schedule(dynamic, 16)
    for (int k = 0; k < o->n; k++) {
      Particle* p = &o->ps[k];
      if (!ifDem(p)) continue;
      calcNeighboursCoupling(o, p, &nb);
      calcFPI(o, p, &nb);
    }
    free(nb.e);
  }
}
/* calcDEMWallContacts sph.h:1634-1674: the nearest boundary or sphBoundary
 * particle in the DEM's neighbour list.  The list of a DEM particle comes from
 * the DEM cells, which hold DEM particles only (DEM are binned in dem_cells
 * alone, dataLL.cpp:979-990), so the scan never finds a wall particle and the
 * force is never applied; restated as the scan. */
static void calcDEMWallContacts(Oracle* o, Particle* p) {
  (void)o;
  int found = 0;
  for (int i = 0; i < p->nnb; i++) {
    const Particle* pn = p->nb[i];
    if ((pn->iam != IAM_BOUNDARY) && (pn->iam != IAM_SPHB)) continue;
    found = 1;
  }
  if (found) {
    fprintf(stderr, "oracle: a wall particle in a DEM neighbour list (dem cells hold DEM only)\n");
    abort();
  }
}
static void passDemWallContacts(Oracle* o) { /* neighboursGroup<calcDEMWallContacts,ifDem> */
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifDem(p)) continue;
    ensureNeighbours(o, p);
    calcDEMWallContacts(o, p);
  }
}
/* DEM_BLOCK sphIncompress.cpp:231-238: calcAvForce then setAvForce
 * (sphIncompress.h:107-114) over the DEM of the container, in its order */
static void demBlock(Oracle* o) {
  double av[3] = {0, 0, 0}, tmp = 0;
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifDem(p)) continue;
    for (int i = 0; i < 3; i++) av[i] += p->f[i];
    tmp++;
  }
  if (tmp == 0) return;
  for (int i = 0; i < 3; i++) av[i] = av[i] / tmp;
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifDem(p)) continue;
    for (int i = 0; i < 3; i++) p->f[i] = av[i];
  }
}
static void passDemContacts(Oracle* o) { /* neighboursGroup<calcDEMContacts,ifDem> */
#pragma omp parallel for schedule(dynamic, 64)
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifDem(p)) continue;
    ensureNeighbours(o, p);
    calcDEMContacts(o, p);
  }
}

/* --------------------------------------------------- CsphIncompress steps */
/* CsphIncompress::start sphIncompress.cpp:3-38 */
static int stepStart(Oracle* o) {
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) driftRAndKick(o, &o->ps[k]);
  if (isLiqDem(o))
    for (int k = 0; k < o->n; k++)
      if (ifDem(&o->ps[k])) driftR(o, &o->ps[k]);
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) driftRest(o, &o->ps[k]);
  o->time += o->dt / 2;
  if ((o->sphStep % o->P.reinit_every == 0) || (o->sphStep == o->P.reinit_at)) {
    if (resetO(o)) return -1;
    for (int i = 0; i < 10; i++) passDensityOnce(o);
  }
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) initSumsMiddle(&o->ps[k]);
  return 0;
}
/* CsphIncompress::middle sphIncompress.cpp:40-165 */
static int stepMiddle(Oracle* o) {
  o->newDt = o->maxdt;
  if (resetO(o)) return -1;
  if (isLiqDem(o)) passPorosity(o);
  if (o->P.flags & SPHDEM_F_VAR_H_CORRECTION) {
    passOmega(o);
    syncOmega(o);
  }
  passEos(o);
  passForce(o);
  if (isLiqDem(o)) {
    o->newDt = dmin2(o->newDt, o->P.liq_dem_timestep);
    if (!(o->P.flags & (SPHDEM_F_FIXED_DEM | SPHDEM_F_NO_DEM_CONTACTS)))
      if (o->time > o->P.time_drop) o->newDt = dmin2(o->newDt, o->P.dem_timestep);
  }
  double newDt = o->newDt;
#pragma omp parallel for schedule(static) reduction(min : newDt)
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (ifSphOrSphBoundary(p) || ifDem(p)) accelCondition(p, &newDt);
  }
  o->newDt = newDt;
  double oldDt = o->dt;
  o->dt = o->newDt; /* both branches of sphIncompress.cpp:145-152 */
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) kick(o, &o->ps[k], 0.5 * oldDt, 0.5 * o->dt);
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) initSumsEnd(&o->ps[k]);
  if (isLiqDem(o)) {
    for (int k = 0; k < o->n; k++)
      if (ifDem(&o->ps[k])) driftRAndKick(o, &o->ps[k]);
    for (int k = 0; k < o->n; k++)
      if (ifDem(&o->ps[k])) initSumsMiddle(&o->ps[k]);
  }
  return 0;
}
/* CsphIncompress::end sphIncompress.cpp:167-251 */
static int stepEnd(Oracle* o) {
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) driftR(o, &o->ps[k]);
  o->time += o->dt / 2;
  if (resetO(o)) return -1;
  if (o->P.flags & SPHDEM_F_VAR_H_CORRECTION) passOmega(o);
  passDddt(o);
  if (isLiqDem(o)) {
    passFpi(o);
    if (!(o->P.flags & (SPHDEM_F_FIXED_DEM | SPHDEM_F_NO_DEM_CONTACTS))) {
      passDemContacts(o);
      if (!(o->P.flags & SPHDEM_F_CUSTOM_WALL)) passDemWallContacts(o);
    }
    if (o->P.flags & SPHDEM_F_DEM_BLOCK) demBlock(o);
    for (int k = 0; k < o->n; k++)
      if (ifDem(&o->ps[k])) kick(o, &o->ps[k], 0.5 * o->dt, 0.5 * o->dt);
  }
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) driftRest(o, &o->ps[k]);
  for (int k = 0; k < o->n; k++)
    if (ifSphOrSphBoundary(&o->ps[k])) calcEnergies(o, &o->ps[k]);
  return 0;
}

/* --------------------------------------------------- sediment3D hooks */
static int typeSelected(uint32_t mask, int iam) { return iam < 32 && ((mask >> iam) & 1u); }
static void hookBodyForce(Oracle* o, uint32_t mask, const double* a) {
  for (int k = 0; k < o->n; k++)
    if (typeSelected(mask, o->ps[k].iam))
      for (int i = 0; i < 3; i++)
        if (a[i] != 0.0) o->ps[k].f[i] += a[i];
}
/* addCustomBoundaries sediment3D/customSim.cpp:67-81 */
static void hookDemFloor(Oracle* o) {
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!ifDem(p)) continue;
    const double overlap = o->P.rmin[2] + o->P.dem_radius - p->r[2];
    if (overlap > 0.0) {
      const double normal[3] = {0.0, 0.0, 1.0};
      const double overlap_dot = -p->vhat[2];
      const double s = demNormalScale(o, overlap, overlap_dot);
      for (int i = 0; i < 3; i++) {
        const double fN = s * normal[i] / p->mass;
        p->fb[i] += fN;
        p->f[i] += fN;
      }
    }
  }
}
/* freezeParticles sediment3D/customSim.cpp:12-25 */
static void hookFreeze(Oracle* o, uint32_t mask, double tfreeze) {
  if (!(o->time <= tfreeze)) return;
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (!typeSelected(mask, p->iam)) continue;
    for (int i = 0; i < 3; i++) {
      p->f[i] = 0.0;
      p->v[i] = 0.0;
      p->vhat[i] = p->v[i];
    }
  }
}
/* sumDragForces sediment3D/customSim.cpp:27-42 */
static void hookSumDrag(Oracle* o) {
  for (int i = 0; i < 6; i++) o->custom[i] = 0.0;
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    if (p->iam == IAM_SPH)
      for (int i = 0; i < 3; i++) o->custom[i] += p->mass * p->fdrag[i];
    else if (p->iam == IAM_DEM)
      for (int i = 3; i < 6; i++) o->custom[i] += p->mass * p->fdrag[i - 3];
  }
}

/* ------------------------------------------------ output-step diagnostics */
/* CsphIncompress::calcOutputVars sphIncompress.h:271-319 with calcOutput
 * (:233-260), calcVarDens (:262-264) and calcAveDensFromMass (:266-272), at
 * one rank (sumOverProcs is the identity).  Sums run in container order. */
static void outputVars(Oracle* o, sphdem_output* g) {
  memset(g, 0, sizeof(*g));
  double nSph = 0, rmsFF = 0, aveDensFromMass = 0;
  for (int k = 0; k < o->n; k++) { /* traverse<calcOutput,ifSphOrSphBoundary> */
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    if (p->iam == IAM_SPH) {
      for (int i = 0; i < 3; i++) g->linMom[i] += p->v[i] * p->mass;
      g->angMom += (p->r[0] * p->v[1] - p->r[1] * p->v[0]) * p->mass;
      g->eKE += 0.5 * p->mass * len2v(p->vhat);
      g->eViscF += p->mass * p->eViscF;
      g->edViscFdt += p->mass * p->deViscFdt;
      g->eViscB += p->mass * p->eViscB;
      g->eBForce += p->mass * p->eBForce;
      g->eFF += p->mass * p->eFF;
      g->edFFdt += p->mass * p->deFFdt;
      nSph++;
    }
    const double v = lenv(p->v);
    if (v > g->maxV) g->maxV = v;
    const double f = lenv(p->f);
    if (f > g->maxF) g->maxF = f;
    const double ff = lenv(p->ff);
    if (ff > g->maxFF) g->maxFF = ff;
    rmsFF += ff * ff; /* len2(ff) of the scalar ff: its square */
    g->aveDens += p->dens;
    g->eElast += p->mass * p->u;
    g->edElastdt += p->mass * p->dudt;
    double denfac = p->dens / o->P.refd;
    denfac = o->P.prb * ((pow(denfac, 6) - 1) / 6 + 1 / denfac - 1);
    g->eElastExact += p->mass * denfac / o->P.refd;
  }
  for (int k = 0; k < o->n; k++) { /* neighboursGroup<calcAveDensFromMass,...> */
    Particle* p = &o->ps[k];
    if (!ifSphOrSphBoundary(p)) continue;
    ensureNeighbours(o, p);
    const double oldDens = p->dens;
    calcDensity(o, p);
    aveDensFromMass += pow(p->dens - oldDens, 2);
    p->dens = oldDens;
  }
  const int n = o->n; /* getParticles()->size(): deleted particles are gone */
  g->n = n;
  g->nSph = (int64_t)nSph;
  g->aveDens = g->aveDens / n;
  g->aveDensFromMass = sqrt(aveDensFromMass / n);
  g->rmsFF = sqrt(rmsFF / n);
  for (int k = 0; k < o->n; k++) { /* traverse<calcVarDens,ifSphOrSphBoundary> */
    Particle* p = &o->ps[k];
    if (ifSphOrSphBoundary(p)) g->varDens += pow(p->dens - g->aveDens, 2);
  }
  g->varDens /= n;
  g->eTotal = g->eKE + g->eElast + g->eViscF + g->eViscB + g->eBForce + g->eFF;
}

/* ============================================================ public API */
#define OEXPORT __attribute__((visibility("default")))

OEXPORT void* orc_create(const sphdem_params* P, int n, const double* r, const double* v,
                         const double* vhat, const double* mass, const double* h,
                         const double* dens, const int* iam) {
  Oracle* o = (Oracle*)calloc(1, sizeof(Oracle));
  o->P = *P;
  o->cap = n + 16;
  o->ps = (Particle*)calloc(o->cap, sizeof(Particle));
  for (int k = 0; k < n; k++) {
    Particle* p = &o->ps[k];
    p->porosity = 1.0; /* Cparticle() particle.h:80 */
    for (int i = 0; i < 3; i++) {
      p->r[i] = r[3 * k + i];
      p->v[i] = v ? v[3 * k + i] : 0.0;
      p->vhat[i] = vhat ? vhat[3 * k + i] : 0.0;
    }
    p->mass = mass[k];
    p->h = h[k];
    p->dens = dens[k];
    p->iam = iam[k];
    p->tag = k + 1;
  }
  o->n = n;
  o->maxdt = P->maxdt;
  /* CdataLL::CdataLL dataLL.cpp:8-40 */
  o->hmax = P->h;
  o->ssr = P->h;
  resizeCells(o);
  if (P->flags & SPHDEM_F_LIQ_DEM) resizeDemCells(o);
  for (int d = 0; d < 27; d++) o->procGhostMax2H[d] = 1.1 * KERNAL_RADIUS * P->h;
  if (resetO(o)) {
    free(o->ps);
    free(o);
    return NULL;
  }
  return o;
}

OEXPORT void orc_destroy(void* vo) {
  Oracle* o = (Oracle*)vo;
  if (!o) return;
  for (int k = 0; k < o->cap; k++) free(o->ps[k].nb);
  for (int k = 0; k < o->ng; k++) free(o->gh[k].nb);
  size_t nc = (size_t)o->cdim[0] * o->cdim[1] * o->cdim[2];
  for (size_t i = 0; i < nc; i++) free(o->cells[i].e);
  free(o->cells);
  if (o->dem_cells) {
    nc = (size_t)o->ddim[0] * o->ddim[1] * o->ddim[2];
    for (size_t i = 0; i < nc; i++) free(o->dem_cells[i].e);
    free(o->dem_cells);
  }
  free(o->ps);
  free(o->gh);
  free(o->ghost_origin);
  free(o->ghost_dir);
  free(o);
}

OEXPORT void orc_set_params(void* vo, const sphdem_params* P) { ((Oracle*)vo)->P = *P; }

/* Host threads of the parallel gathers (1 = the serial sweep); returns the
 * number in use afterwards. */
OEXPORT void orc_output_vars(void* vo, sphdem_output* out) { outputVars((Oracle*)vo, out); }

OEXPORT int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

OEXPORT int orc_num_particles(void* vo) { return ((Oracle*)vo)->n; }
OEXPORT long orc_num_deleted(void* vo) { return ((Oracle*)vo)->ndeleted; }
OEXPORT int orc_num_ghosts(void* vo) { return ((Oracle*)vo)->ng; }
OEXPORT double orc_search_radius(void* vo) { return ((Oracle*)vo)->ssr; }

OEXPORT void orc_set_globals(void* vo, const sphdem_globals* g) {
  Oracle* o = (Oracle*)vo;
  o->time = g->time;
  o->dt = g->dt;
  o->newDt = g->newDt;
  o->maxdt = g->maxdt;
  o->sphStep = (long)g->sphStep;
  for (int i = 0; i < 9; i++) o->custom[i] = g->custom[i];
}
OEXPORT void orc_get_globals(void* vo, sphdem_globals* g) {
  Oracle* o = (Oracle*)vo;
  g->time = o->time;
  g->dt = o->dt;
  g->newDt = o->newDt;
  g->maxdt = o->maxdt;
  g->sphStep = o->sphStep;
  for (int i = 0; i < 9; i++) g->custom[i] = o->custom[i];
}

static double* fieldPtr(Particle* p, int field, int* width) {
  *width = 3;
  switch (field) {
    case SPHDEM_FIELD_R: return p->r;
    case SPHDEM_FIELD_V: return p->v;
    case SPHDEM_FIELD_V0: return p->v0;
    case SPHDEM_FIELD_VHAT: return p->vhat;
    case SPHDEM_FIELD_VHAT0: return p->vhat0;
    case SPHDEM_FIELD_F: return p->f;
    case SPHDEM_FIELD_FP: return p->fp;
    case SPHDEM_FIELD_FV: return p->fv;
    case SPHDEM_FIELD_FB: return p->fb;
    case SPHDEM_FIELD_FDRAG: return p->fdrag;
    case SPHDEM_FIELD_FVEL: return p->fvel;
    default: break;
  }
  *width = 1;
  switch (field) {
    case SPHDEM_FIELD_DENS: return &p->dens;
    case SPHDEM_FIELD_DDDT: return &p->dddt;
    case SPHDEM_FIELD_H: return &p->h;
    case SPHDEM_FIELD_MASS: return &p->mass;
    case SPHDEM_FIELD_PRESS: return &p->press;
    case SPHDEM_FIELD_PDR2: return &p->pdr2;
    case SPHDEM_FIELD_OMEGA: return &p->omega;
    case SPHDEM_FIELD_POROSITY: return &p->porosity;
    case SPHDEM_FIELD_SHEPSUM: return &p->shepSum;
    case SPHDEM_FIELD_U: return &p->u;
    case SPHDEM_FIELD_DUDT: return &p->dudt;
    default: break;
  }
  *width = 0;
  return NULL;
}

/* Field values indexed by original index (tag - 1); deleted particles keep
 * whatever the caller pre-filled. */
OEXPORT int orc_get_field(void* vo, int field, void* out) {
  Oracle* o = (Oracle*)vo;
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    int idx = p->tag - 1;
    if (field == SPHDEM_FIELD_IAM) {
      ((int32_t*)out)[idx] = p->iam;
      continue;
    }
    if (field == SPHDEM_FIELD_TAG) {
      ((int32_t*)out)[idx] = p->tag;
      continue;
    }
    int w;
    double* src = fieldPtr(p, field, &w);
    if (!src) return -1;
    for (int i = 0; i < w; i++) ((double*)out)[(size_t)w * idx + i] = src[i];
  }
  return 0;
}
OEXPORT int orc_set_field(void* vo, int field, const void* in) {
  Oracle* o = (Oracle*)vo;
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    int idx = p->tag - 1;
    if (field == SPHDEM_FIELD_IAM) {
      p->iam = ((const int32_t*)in)[idx];
      continue;
    }
    int w;
    double* dst = fieldPtr(p, field, &w);
    if (!dst) return -1;
    for (int i = 0; i < w; i++) dst[i] = ((const double*)in)[(size_t)w * idx + i];
  }
  /* ghosts mirror their origin for the fields the reference ships with them */
  return 0;
}
/* Refresh ghost copies of every synced field from their origins (what the
 * reference's reset + sync calls would have produced for the new values). */
OEXPORT void orc_refresh_ghosts(void* vo) {
  Oracle* o = (Oracle*)vo;
  for (int g = 0; g < o->ng; g++) {
    Particle* src = o->ghost_origin[g];
    Particle* gg = &o->gh[g];
    int shift[3] = {gg->shift[0], gg->shift[1], gg->shift[2]};
    double r[3] = {gg->r[0], gg->r[1], gg->r[2]};
    Particle** nb = gg->nb;
    int capnb = gg->capnb;
    *gg = *src;
    gg->nb = nb;
    gg->capnb = capnb;
    gg->nnb = 0;
    for (int i = 0; i < 3; i++) {
      gg->r[i] = r[i];
      gg->shift[i] = shift[i];
    }
  }
}

OEXPORT int orc_reset(void* vo) { return resetO((Oracle*)vo); }
OEXPORT int orc_start(void* vo) { return stepStart((Oracle*)vo); }
OEXPORT int orc_middle(void* vo) { return stepMiddle((Oracle*)vo); }
OEXPORT int orc_end(void* vo) { return stepEnd((Oracle*)vo); }

OEXPORT void orc_pass_omega(void* vo) {
  passOmega((Oracle*)vo);
  syncOmega((Oracle*)vo);
}
OEXPORT void orc_pass_eos(void* vo) { passEos((Oracle*)vo); }
OEXPORT void orc_pass_force(void* vo) { passForce((Oracle*)vo); }
OEXPORT void orc_pass_dddt(void* vo) { passDddt((Oracle*)vo); }
OEXPORT void orc_pass_density(void* vo) { passDensityOnce((Oracle*)vo); }
OEXPORT void orc_pass_porosity(void* vo) { passPorosity((Oracle*)vo); }
OEXPORT void orc_pass_fpi(void* vo) { passFpi((Oracle*)vo); }
OEXPORT void orc_pass_dem_contacts(void* vo) { passDemContacts((Oracle*)vo); }
OEXPORT double orc_new_dt(void* vo) { return ((Oracle*)vo)->newDt; }
OEXPORT void orc_set_new_dt(void* vo, double v) { ((Oracle*)vo)->newDt = v; }

OEXPORT void orc_hook_body_force(void* vo, uint32_t mask, const double* a) {
  hookBodyForce((Oracle*)vo, mask, a);
}
OEXPORT void orc_hook_dem_floor(void* vo) { hookDemFloor((Oracle*)vo); }
OEXPORT void orc_hook_freeze(void* vo, uint32_t mask, double t) { hookFreeze((Oracle*)vo, mask, t); }
OEXPORT void orc_hook_sum_drag(void* vo) { hookSumDrag((Oracle*)vo); }

/* run.cpp:102-118 with a hook set (SPHDEM_HOOKS_*) */
OEXPORT int orc_run_steps(void* vo, int nsteps, int hookset) {
  Oracle* o = (Oracle*)vo;
  const double g[3] = {0.0, 0.0, -9.81};
  for (int s = 0; s < nsteps; s++) {
    o->sphStep++;
    if (stepStart(o)) return -1;
    if (hookset == SPHDEM_HOOKS_DAMBREAK || hookset == SPHDEM_HOOKS_SEDIMENT)
      hookBodyForce(o, 1u << IAM_SPH, g); /* beforeMiddle */
    if (stepMiddle(o)) return -1;
    if (hookset == SPHDEM_HOOKS_SEDIMENT) { /* beforeEnd sediment3D/customSim.cpp:97-107 */
      hookSumDrag(o);
      hookBodyForce(o, 1u << IAM_DEM, g);
      hookDemFloor(o);
      hookFreeze(o, 1u << IAM_DEM, o->P.time_drop);
    }
    if (stepEnd(o)) return -1;
    if (hookset == SPHDEM_HOOKS_SEDIMENT) hookFreeze(o, 1u << IAM_DEM, o->P.time_drop); /* afterEnd */
  }
  return 0;
}

/* Cell coordinates (getCellI / dem_getCellI) per original index. */
OEXPORT void orc_cells(void* vo, int32_t* ijk) {
  Oracle* o = (Oracle*)vo;
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    int c[3];
    if ((o->P.flags & SPHDEM_F_LIQ_DEM) && p->iam == IAM_DEM)
      demGetCellI(o, p->r, c);
    else
      getCellI(o, p->r, c);
    for (int i = 0; i < 3; i++) ijk[3 * (p->tag - 1) + i] = c[i];
  }
}

static int imageCode(const Particle* b) { return (b->shift[0] + 1) + 3 * (b->shift[1] + 1) + 9 * (b->shift[2] + 1); }

/* The pair set of a search kind as (tag-1, tag-1, image) triples.  For the SPH
 * and DEM kinds the lists are the cached ones (built now if empty). */
OEXPORT long orc_pairs(void* vo, int kind, int32_t* a, int32_t* b, int32_t* img, long cap) {
  Oracle* o = (Oracle*)vo;
  long cnt = 0;
  PList nb = {0, 0, 0};
  for (int k = 0; k < o->n; k++) {
    Particle* p = &o->ps[k];
    Particle** list = NULL;
    int nl = 0;
    if (kind == SPHDEM_PAIRS_SPH) {
      if (!ifSphOrSphBoundary(p)) continue;
      ensureNeighbours(o, p);
      list = p->nb;
      nl = p->nnb;
    } else if (kind == SPHDEM_PAIRS_DEM) {
      if (!ifDem(p)) continue;
      ensureNeighbours(o, p);
      list = p->nb;
      nl = p->nnb;
    } else if (kind == SPHDEM_PAIRS_SPH2DEM) {
      if (!ifSphOrSphBoundary(p)) continue;
      calcNeighboursCoupling(o, p, &nb);
      list = nb.e;
      nl = nb.n;
    } else {
      if (!ifDem(p)) continue;
      calcNeighboursCoupling(o, p, &nb);
      list = nb.e;
      nl = nb.n;
    }
    for (int i = 0; i < nl; i++) {
      if (a && cnt < cap) {
        a[cnt] = p->tag - 1;
        b[cnt] = list[i]->tag - 1;
        img[cnt] = imageCode(list[i]);
      }
      cnt++;
    }
  }
  free(nb.e);
  return cnt;
}

/* Kernel family exposed for the closed-form tests. */
OEXPORT double orc_kernel_W(void* vo, double q, double h) { return kW((Oracle*)vo, q, h); }
OEXPORT double orc_kernel_F(void* vo, double q, double h) { return kF((Oracle*)vo, q, h); }
OEXPORT double orc_kernel_K(void* vo, double q, double h) { return kK((Oracle*)vo, q, h); }
OEXPORT double orc_kernel_pdWh(void* vo, double q, double h) { return kpdWh((Oracle*)vo, q, h); }
OEXPORT void orc_particle_drag(void* vo, double porosity, const double* vhat, const double* fvel,
                               double mass, double* out) {
  Particle p;
  memset(&p, 0, sizeof(p));
  p.porosity = porosity;
  p.mass = mass;
  for (int i = 0; i < 3; i++) p.vhat[i] = vhat[i];
  calcParticleDrag((Oracle*)vo, &p, fvel, out);
}
