// sphdem_kernels.hip -- gfx950 kernels of the SPH-DEM per-timestep loop, FP64 form.
//
// FP64 kernels are compiled without FMA contraction so that every a*b+c rounds
// twice, as the reference's x86-64 build does; operand order follows the
// reference expression by expression (cited per kernel).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/sphdem_c.h"
#include "sphdem_kernels.h"
#include "sphdem_math.h"

namespace sphdem {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ bool isSphT(int t) { return t == T_SPH || t == T_SPHB; }
__device__ __forceinline__ double imgc(double r, int s, double Lp, double Lm) {
  return s == 1 ? (Lp + r) : (s == -1 ? (Lm + r) : r);
}
__device__ __forceinline__ int cellOf(double x, double cmin, double edge, double off) {
  return (int)((x - cmin) / edge + off);
}
// (int)((x - o) / e + add), the reference's cell arithmetic, with the
// quotient first taken as the product by inv (1/e rounded): that differs from
// the quotient by a few units in the last place, so both truncate alike
// unless the value lies within 1e-9 of an integer, and only then is the
// division done.  The same cell as the division for every x.
__device__ __forceinline__ int cellTrunc(double x, double o, double e, double inv, double add) {
  const double t = (x - o) * inv + add;
  if (fabs(t - rint(t)) > 1e-9) return (int)t;
  return (int)((x - o) / e + add);
}
__device__ __forceinline__ int cellOf(double x, const Grid& g, int d) {
  return cellTrunc(x, g.cmin[d], g.edge, g.inv, g.off);
}
__device__ __forceinline__ double dmin2(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax2(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}
__device__ __forceinline__ double& comp(double4& v, int k) { return (&v.x)[k]; }

// Wave (64-lane) and block reductions.
__device__ __forceinline__ double waveMin(double v) {
  for (int d = 32; d > 0; d >>= 1) v = dmin2(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ double waveMax(double v) {
  for (int d = 32; d > 0; d >>= 1) v = dmax2(v, __shfl_xor(v, d, 64));
  return v;
}
// Deterministic wavefront sum: a fixed butterfly, so the result does not
// depend on timing (every lane ends with the same value).
__device__ __forceinline__ double waveSum(double v) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
// positive doubles order like their bit patterns, so min/max go through the
// 64-bit integer atomics and are exact and order independent; one atomic per
// block (blockDim.x <= 1024)
__device__ __forceinline__ void atomicMinPos(unsigned long long* p, double v) {
  __shared__ double red[16];
  v = waveMin(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) v = dmin2(v, red[w]);
    atomicMin(p, dbits(v));
  }
}
__device__ __forceinline__ void atomicMaxPos(unsigned long long* p, double v) {
  __shared__ double red[16];
  v = waveMax(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) v = dmax2(v, red[w]);
    atomicMax(p, dbits(v));
  }
}

// Visits every particle of grid g in the cells within `reach` of the query
// point q, for each periodic image shift s (s = 0 first): the candidates b are
// real particles whose image r_b + s*L may lie near q.  Cells along z of one
// (x,y) column are contiguous in the sorted store, so each column is one range.
// `first`/`stride` split each contiguous range over the lanes of a wavefront.
struct KeepAll {
  __device__ bool operator()(int, int, int) const { return true; }
};
template <class V, class K = KeepAll>
__device__ __forceinline__ void forCells(const Grid& g, const int* __restrict__ start,
                                         const Box& B, double qx, double qy, double qz,
                                         int reach, V&& visit, int first = 0, int stride = 1,
                                         K&& keep = K()) {
  const int nsx = B.periodic[0] ? 3 : 1, nsy = B.periodic[1] ? 3 : 1, nsz = B.periodic[2] ? 3 : 1;
  const double q[3] = {qx, qy, qz};
  for (int ix = 0; ix < nsx; ix++) {
    const int sx = ix == 0 ? 0 : (ix == 1 ? -1 : 1);
    const double tx = sx == 0 ? q[0] : (sx == 1 ? q[0] - B.Lp[0] : q[0] - B.Lm[0]);
    const int cx = cellOf(tx, g, 0);
    const int x0 = max(cx - reach, 0), x1 = min(cx + reach, g.dim[0] - 1);
    if (x0 > x1) continue;
    for (int iy = 0; iy < nsy; iy++) {
      const int sy = iy == 0 ? 0 : (iy == 1 ? -1 : 1);
      const double ty = sy == 0 ? q[1] : (sy == 1 ? q[1] - B.Lp[1] : q[1] - B.Lm[1]);
      const int cy = cellOf(ty, g, 1);
      const int y0 = max(cy - reach, 0), y1 = min(cy + reach, g.dim[1] - 1);
      if (y0 > y1) continue;
      for (int iz = 0; iz < nsz; iz++) {
        const int sz = iz == 0 ? 0 : (iz == 1 ? -1 : 1);
        const double tz = sz == 0 ? q[2] : (sz == 1 ? q[2] - B.Lp[2] : q[2] - B.Lm[2]);
        const int cz = cellOf(tz, g, 2);
        const int z0 = max(cz - reach, 0), z1 = min(cz + reach, g.dim[2] - 1);
        if (z0 > z1 || !keep(cx, cy, cz)) continue;
        for (int a = x0; a <= x1; a++)
          for (int b = y0; b <= y1; b++) {
            const int k0 = g.base + (a * g.dim[1] + b) * g.dim[2];
            const int j0 = start[k0 + z0], j1 = start[k0 + z1 + 1];
            for (int j = j0 + first; j < j1; j += stride) visit(j, sx, sy, sz);
          }
      }
    }
  }
}

struct Img {
  double x, y, z;
};
__device__ __forceinline__ Img imagePos(const double4& p, int sx, int sy, int sz, const Box& B) {
  return Img{imgc(p.x, sx, B.Lp[0], B.Lm[0]), imgc(p.y, sy, B.Lp[1], B.Lm[1]),
             imgc(p.z, sz, B.Lp[2], B.Lm[2])};
}

// gradW (sph.h:322-342): dr*F(q,h), divided by pa.omega under VAR_H_CORRECTION
__device__ __forceinline__ void gradW(double wcon, bool varh, double omega_a, double dx,
                                      double dy, double dz, double q, double h, double* g) {
  double drF0 = 0.0, drF1 = 0.0, drF2 = 0.0;
  if (q != 0) {
    const double F = wendF(wcon, q, h);
    drF0 = dx * F;
    drF1 = dy * F;
    drF2 = dz * F;
  }
  if (varh) {
    g[0] = drF0 / omega_a;
    g[1] = drF1 / omega_a;
    g[2] = drF2 / omega_a;
  } else {
    g[0] = drF0;
    g[1] = drF1;
    g[2] = drF2;
  }
}

// Which periodic shifts of a particle along axis d give a ghost inside the band.
__device__ __forceinline__ int ghostShiftsAxis(double x, int d, const Grid& g, const Box& B,
                                               const GhostBand& gb, int* cell) {
  // bit 0: shift -1 valid, bit 1: shift +1 valid; cell[0]/cell[1] their cells
  int m = 0;
  if (!B.periodic[d]) return 0;
  const int lo = gb.lo[d], hi = gb.hi[d];
  const int cm = cellOf(B.Lm[d] + x, g, d);
  const int cp = cellOf(B.Lp[d] + x, g, d);
  if (cm >= lo && cm <= hi) m |= 1;
  if (cp >= lo && cp <= hi) m |= 2;
  cell[0] = cm;
  cell[1] = cp;
  return m;
}
__device__ __forceinline__ int popc2(int m) { return (m & 1) + ((m >> 1) & 1); }

__device__ __forceinline__ int ghostCount(const double4& p, const Grid& g, const Box& B,
                                          const GhostBand& gb) {
  int cc[2];
  const int mx = ghostShiftsAxis(p.x, 0, g, B, gb, cc);
  const int my = ghostShiftsAxis(p.y, 1, g, B, gb, cc);
  const int mz = ghostShiftsAxis(p.z, 2, g, B, gb, cc);
  return (1 + popc2(mx)) * (1 + popc2(my)) * (1 + popc2(mz)) - 1;
}
// ============================================================ integrator parts
// start() per particle (sphIncompress.cpp:4-16; sphIncompress.h:620-704,
// 744-763) in two parts, so that the reset that follows it in a step can do
// them inside its own passes with the same arithmetic:
//  - startDrift: r += dt vhat (sph, sphBoundary, dem), then driftRest's
//    dens += dt dddt and h = HFAC (mass/dens)^(1/3) (sph, sphBoundary);
//  - startKick (sph, sphBoundary): v0 = v and vhat0 = vhat, v and vhat
//    kicked by dt f (sph only), then initSumsMiddle (f, fp.xyz, fv, fb = 0)
//    and the beforeMiddle body force for the types in bodyMask.
// Neither reads a field the other writes.
__device__ __forceinline__ void startDrift(const Store& S, int a, int t, unsigned flags,
                                           double hfac, double dt, double4& p) {
  const double4 vh = S.vhm[a];
  p.x += dt * vh.x;
  p.y += dt * vh.y;
  p.z += dt * vh.z;
  if (isSphT(t)) {
    double4 d = S.dpo[a];
    d.x += dt * S.vh0d[a].w;
    S.dpo[a] = d;
    if (!(flags & SPHDEM_F_CONST_H)) p.w = hfac * pow_third(vh.w / d.x);
  }
}
__device__ __forceinline__ void halfKick(double4& v, const double4& f, double dt) {
  v.x += dt * f.x;
  v.y += dt * f.y;
  v.z += dt * f.z;
}
// src entry i to dst entry q (the same entry in place); MOVE: also copy the
// fields the kick leaves alone
template <bool MOVE>
__device__ __forceinline__ void startKick(const Store& src, int i, const Store& dst, int q, int t,
                                          double dt, unsigned bodyMask, double ax, double ay,
                                          double az) {
  if (t != T_SPHB) {
    double4 v = src.vel[i];
    double4 vh = src.vhm[i];
    const double4 f = src.frc[i];
    dst.v0p[q] = make_double4(v.x, v.y, v.z, src.v0p[i].w);
    dst.vh0d[q] = make_double4(vh.x, vh.y, vh.z, src.vh0d[i].w);
    halfKick(v, f, dt);
    halfKick(vh, f, dt);
    dst.vel[q] = v;
    dst.vhm[q] = vh;
  } else if (MOVE) {
    dst.vel[q] = src.vel[i];
    dst.vhm[q] = src.vhm[i];
    dst.v0p[q] = src.v0p[i];
    dst.vh0d[q] = src.vh0d[i];
  }
  const double4 z = make_double4(0.0, 0.0, 0.0, 0.0);
  double4 f0 = z;
  if ((bodyMask >> t) & 1u) {
    if (ax != 0.0) f0.x += ax;
    if (ay != 0.0) f0.y += ay;
    if (az != 0.0) f0.z += az;
  }
  dst.frc[q] = f0;
  dst.fp[q] = make_double4(0.0, 0.0, 0.0, src.fp[i].w);
  dst.fv[q] = z;
  dst.fb[q] = z;
}

// ============================================================ cell build
// CdataLL::updateDomain's migration for one rank whose neighbours are itself
// through periodic faces (dataLL.cpp:370-403, 459-497): wrap or delete, then
// hmax (calcDomainLimits, dataLL.cpp:954-975).
// With gcnt set, also each particle's ghost count on the current grid (used
// when the reset keeps that grid).
// F.mode 1: the reset right after start() does start()'s drift part first
// (its kick part goes into k_permute<true>; a particle deleted here gets both
// parts here); F.mode 2: the reset right after end()'s drift does that drift
// first (driftR on sph, sphIncompress.cpp:170).  Both add dt/2 to the time.
__global__ void k_wrap(Store S, int n, double dmin0, double dmin1, double dmin2_,
                       double dsp0, double dsp1, double dsp2, Box B, DevGlobals* g, Grid gs,
                       GhostBand gb, int* __restrict__ gcnt, int initNewDt, WrapFold F) {
  // middle() starts its dt minimum at maxdt (nothing reads it during the reset)
  if (initNewDt && blockIdx.x == 0 && threadIdx.x == 0)
    g->newDt_bits = (unsigned long long)__double_as_longlong(g->maxdt);
  const double dt = F.mode ? g->dt / 2 : 0.0;
  if (F.mode && blockIdx.x == 0 && threadIdx.x == 0) g->time += g->dt / 2;
  // grid-stride over the particles: a few hundred blocks, so few atomics on hmax
  double hv = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int ghosts = 0;
    const int t = S.iam[i];
    if (t != T_DELETED) {
      double4 p = S.posh[i];
      bool moved = false;
      if (F.mode == 1 && i < F.nlive && (isSphT(t) || t == T_DEM)) {
        startDrift(S, i, t, F.flags, F.hfac, dt, p);
        moved = true;
      } else if (F.mode == 2 && i < F.nlive && isSphT(t)) {
        const double4 vh = S.vhm[i];
        p.x += dt * vh.x;
        p.y += dt * vh.y;
        p.z += dt * vh.z;
        moved = true;
      }
      const double dmin[3] = {dmin0, dmin1, dmin2_};
      const double dsp[3] = {dsp0, dsp1, dsp2};
      int oc[3];
      bool out = false, nonper = false, notone = false;
      for (int k = 0; k < 3; k++) {
        oc[k] = (int)((comp(p, k) - dmin[k]) / dsp[k] + 1);
        if (oc[k] > 2 || oc[k] < 0) out = true;
        if (oc[k] != 1) {
          notone = true;
          if (!B.periodic[k]) nonper = true;
        }
      }
      if (out || nonper) {
        if (moved) S.posh[i] = p;
        if (F.mode == 1 && moved && isSphT(t))
          startKick<false>(S, i, S, i, t, dt, F.bodyMask, F.ax, F.ay, F.az);
        S.iam[i] = T_DELETED;
        atomicAdd(t == T_DEM ? &g->ndel_dem : &g->ndel_sph, 1ull);
      } else {
        if (notone) {
          for (int k = 0; k < 3; k++) {
            if (oc[k] == 0) comp(p, k) = B.Lp[k] + comp(p, k);
            if (oc[k] == 2) comp(p, k) = B.Lm[k] + comp(p, k);
          }
        }
        if (notone || moved) S.posh[i] = p;
        hv = dmax2(hv, p.w);
        if (gcnt && isSphT(t)) ghosts = ghostCount(p, gs, B, gb);
      }
    }
    if (gcnt) gcnt[i] = ghosts;
  }
  atomicMaxPos(&g->hmax_bits, hv);
}

// Slots within a cell for a wavefront's lanes: each run of equal keys
// (k >= 0) in lane order takes its slots with one atomic; lanes with k < 0
// take none.  Returns the lane's slot (undefined for k < 0).
__device__ __forceinline__ int runSlots(int k, int* __restrict__ counts) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(k, 1, 64);
  const bool head = lane == 0 || prev != k;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int hl = 63 - __builtin_clzll(heads & upto);
  int base = 0;
  if (head && k >= 0) {
    const unsigned long long rest = lane == 63 ? 0ull : (heads >> (lane + 1));
    const int len = rest ? __builtin_ctzll(rest) + 1 : 64 - lane;
    base = atomicAdd(&counts[k], len);
  }
  base = __shfl(base, hl, 64);
  return base + (lane - hl);
}

// Binning of the real entries [0, n) (CdataLL::insert, dataLL.cpp:978-997):
// key (cell) of every entry and a slot within its cell.  The store is sorted
// by cell from the previous reset, so neighbouring lanes mostly share a key:
// each run of equal keys in a wavefront takes its slots with one atomic
// (slots only need to be distinct; the rank kernel orders each cell).  The
// first thread also clears the hmax reduction target for the next reset, whose
// value the host has read by now.
__global__ void k_bin(Store S, int n, Grid gs, Grid gc, int keyDeleted, int* __restrict__ key,
                      int* __restrict__ slot, int* __restrict__ counts, DevGlobals* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) g->hmax_bits = 0ull;
  int k = -1;
  if (i < n) {
    const int t = S.iam[i];
    if (t == T_DELETED) {
      k = keyDeleted;
      S.cell[i] = 0;
    } else {
      const double4 p = S.posh[i];
      const Grid& gr = (t == T_DEM) ? gc : gs;
      int c[3];
      for (int d = 0; d < 3; d++) {
        c[d] = cellOf(comp(const_cast<double4&>(p), d), gr, d);
        if (c[d] < 0 || c[d] >= gr.dim[d]) {
          atomicOr(&g->errflags, 1u);
          c[d] = c[d] < 0 ? 0 : gr.dim[d] - 1;
        }
      }
      k = gr.base + (c[0] * gr.dim[1] + c[1]) * gr.dim[2] + c[2];
      S.cell[i] = packCell(c[0], c[1], c[2]);
    }
    key[i] = k;
  }
  const int sl = runSlots(k, counts);
  if (i < n) slot[i] = sl;
}

// ---- ghosts: the periodic images CdataLL::updateDomain receives
// (dataLL.cpp:631-695).  A real SPH-group particle gets a ghost for every
// combination of shifts (+-L on periodic axes) whose image cell, on each
// shifted axis, lies within one cell of the SPH group's cell range: the only
// ghosts a 27-cell stencil of a real particle can reach.  `ord` numbers the
// combinations in the order the reference receives them (receive slot
// c = shift + 1, x-major: ord = 9 (sx+1) + 3 (sy+1) + (sz+1)).
__global__ void k_ghost_count(Store S, int n, Grid g, Box B, GhostBand gb,
                              int* __restrict__ gcnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  gcnt[i] = isSphT(S.iam[i]) ? ghostCount(S.posh[i], g, B, gb) : 0;
}

// Writes ghost entry e = n + goff[i] + k for each ghost of real particle i:
// its key in the ghost range (gbase + image cell), its slot, origin, `ord`
// and packed image cell.
// Ghost entries n + goff[i] ... of each real particle, one per valid
// combination of shifts in receive order.  The 26 combinations are walked
// wave-uniformly, so neighbouring lanes with the same combination (mostly the
// same ghost cell: the store is sorted by cell) share one atomic per run.
// Ghosts beyond `cap` entries are left out (counted nowhere): a reset that
// cannot read the count back bounds them by the ghost entries it processes.
__global__ void k_ghost_make(Store S, int n, Grid g, Box B, GhostBand gb,
                             const int* __restrict__ goff, int gbase, int* __restrict__ key,
                             int* __restrict__ slot, int* __restrict__ counts,
                             int* __restrict__ gsrc, int* __restrict__ gord,
                             int* __restrict__ gcell, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int cnt = i < n ? goff[i + 1] - goff[i] : 0;
  if (__ballot(cnt > 0) == 0ull) return;  // wave-uniform: no lane has ghosts
  int cx[2] = {0, 0}, cy[2] = {0, 0}, cz[2] = {0, 0};
  int mx = 0, my = 0, mz = 0, rc = 0, e = 0;
  if (cnt > 0) {
    const double4 p = S.posh[i];
    rc = S.cell[i];
    mx = ghostShiftsAxis(p.x, 0, g, B, gb, cx);
    my = ghostShiftsAxis(p.y, 1, g, B, gb, cy);
    mz = ghostShiftsAxis(p.z, 2, g, B, gb, cz);
    e = n + goff[i];
  }
  for (int ord = 0; ord < 27; ord++) {
    if (ord == 13) continue;  // no shift: the particle itself
    const int sx = ord / 9 - 1, sy = (ord / 3) % 3 - 1, sz = ord % 3 - 1;
    const bool on = cnt > 0 && (sx == 0 || ((mx >> (sx > 0)) & 1)) &&
                    (sy == 0 || ((my >> (sy > 0)) & 1)) && (sz == 0 || ((mz >> (sz > 0)) & 1)) &&
                    e - n < cap;
    int k = -1, X = 0, Y = 0, Z = 0;
    if (on) {
      X = sx == 0 ? cellX(rc) : cx[sx > 0];
      Y = sy == 0 ? cellY(rc) : cy[sy > 0];
      Z = sz == 0 ? cellZ(rc) : cz[sz > 0];
      k = gbase + (X * g.dim[1] + Y) * g.dim[2] + Z;
    }
    if (__ballot(on) == 0ull) continue;
    const int sl = runSlots(k, counts);
    if (on) {
      key[e] = k;
      slot[e] = sl;
      gsrc[e - n] = i;
      gord[e - n] = ord;
      gcell[e - n] = packCell(X, Y, Z);
      e++;
    }
  }
}

// Exclusive scan, 4096 elements per 1024-thread block: per-thread serial scan
// of 4 items, 64-lane wavefront scan, then across the 16 wavefronts in LDS.
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// copyLast (optional): also receives out[n - 1]
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tile(const int* __restrict__ in,
                                                            int* __restrict__ out,
                                                            int* __restrict__ tileSums, int n,
                                                            int* copyLast) {
  __shared__ int waveTot[SCAN_THREADS / 64];
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int run = 0;
  for (int k = 0; k < SCAN_ITEMS; k++) {
    const int x = (base + k < n) ? in[base + k] : 0;
    v[k] = run;
    run += x;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = run;
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) waveTot[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    int w = (lane < SCAN_THREADS / 64) ? waveTot[lane] : 0;
    int wi = w;
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(wi, d, 64);
      if (lane >= d) wi += y;
    }
    if (lane < SCAN_THREADS / 64) waveTot[lane] = wi - w;
    if (lane == SCAN_THREADS / 64 - 1 && tileSums) tileSums[blockIdx.x] = wi;
  }
  __syncthreads();
  const int excl = waveTot[wave] + incl - run;
  for (int k = 0; k < SCAN_ITEMS; k++)
    if (base + k < n) {
      out[base + k] = excl + v[k];
      if (copyLast && base + k == n - 1) *copyLast = excl + v[k];
    }
}
// Second pass: block b adds the sum of tiles [0, b) to tile b.  The first
// wavefront reduces those tile sums (each block redoes this small prefix,
// which saves a launch for the scan of the tile sums).
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_add(int* __restrict__ out,
                                                           const int* __restrict__ tileSums,
                                                           int n, int* copyLast) {
  __shared__ int pre;
  const int b = blockIdx.x;
  if (threadIdx.x < 64) {
    int acc = 0;
    for (int t = threadIdx.x; t < b; t += 64) acc += tileSums[t];
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if (threadIdx.x == 0) pre = acc;
  }
  __syncthreads();
  if (b == 0) return;
  const int base = b * SCAN_TILE + threadIdx.x;
  for (int k = 0; k < SCAN_ITEMS; k++) {
    const int i = base + k * SCAN_THREADS;
    if (i < n) {
      const int v = out[i] + pre;
      out[i] = v;
      if (copyLast && i == n - 1) *copyLast = v;
    }
  }
}

// Order inside a cell = the reference's insertion order: real particles in
// container order (original index, dataLL.cpp:226-231); ghosts in the order
// they are received (slot `ord`, then the origin's container order).
__device__ __forceinline__ long long orderKey(int e, int n, const int* __restrict__ orig,
                                              const int* __restrict__ gsrc,
                                              const int* __restrict__ gord) {
  if (e < n) return orig[e];
  return ((long long)gord[e - n] << 32) | (unsigned)orig[gsrc[e - n]];
}
// Entries [0, nreal + live ghosts) of a launch over ncap = nreal + ghost
// capacity entries: the live ghost count is read from the device (ngp) when
// the host has not read it back, else the launch count is exact.
__device__ __forceinline__ int liveEntries(int ncap, int nreal, const int* __restrict__ ngp) {
  return ngp ? nreal + min(*ngp, ncap - nreal) : ncap;
}
// Scatter of the entries into their cells' ranges (any order within a cell),
// with each entry's order key beside it.  Also clears the cell counts
// [0, k1) for the next reset (the scan has consumed them).
__global__ void k_scatter(int ncap, int nreal, const int* __restrict__ key,
                          const int* __restrict__ slot, const int* __restrict__ start,
                          int* __restrict__ tmp, long long* __restrict__ okey,
                          const int* __restrict__ orig, const int* __restrict__ gsrc,
                          const int* __restrict__ gord, int* __restrict__ counts, int k1,
                          const int* __restrict__ ngp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = liveEntries(ncap, nreal, ngp);
  if (i < n) {
    const int q = start[key[i]] + slot[i];
    tmp[q] = i;
    okey[q] = orderKey(i, nreal, orig, gsrc, gord);
  }
  for (int k = i; k < k1; k += gridDim.x * blockDim.x) counts[k] = 0;
}

// Each entry's rank is the number of cell members with a smaller order key.
__global__ void k_rank(int ncap, int nreal, const int* __restrict__ key,
                       const int* __restrict__ start, const int* __restrict__ tmp,
                       const long long* __restrict__ okey, int* __restrict__ perm,
                       const int* __restrict__ ngp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= liveEntries(ncap, nreal, ngp)) return;
  const int i = tmp[p];
  const int k = key[i];
  const int s0 = start[k], s1 = start[k + 1];
  const long long o = okey[p];
  int rank = 0;
  for (int q = s0; q < s1; q++) rank += (okey[q] < o) ? 1 : 0;
  perm[s0 + rank] = i;
}

// Permutation into the second buffer.  A ghost entry copies its origin with
// the position shifted by its periods (the ghost's r = r + (RMAX-RMIN) or
// r + (RMIN-RMAX) per shifted axis, dataLL.cpp:651-666); inv maps the real
// entries' old positions to their new ones.
__device__ __forceinline__ void ordShift(int ord, int& sx, int& sy, int& sz) {
  sx = ord / 9 - 1;
  sy = (ord / 3) % 3 - 1;
  sz = ord % 3 - 1;
}
// The records of PermRecords for store entry p from its new position-and-h
// q and vhat-and-mass vh.
__device__ __forceinline__ void permRecords(const PermRecords& R, int p, const double4& q,
                                            const double4& vh) {
  if (R.lists) {
    uint2 r;
    if (!listRecord(q, R.lq, r)) atomicOr(&R.lflag[R.gen & 1], 1);
    R.rec[p] = r;
  }
  if (R.mixed) {
    R.P[p] = mixRecordP(q, vh, R.Q);
    R.V[p] = mixRecordV(q, vh);
  }
}
// START: the reset right after start(), whose kick part (startKick) is done
// here on the way (k_wrap did its drift part); a ghost takes its origin's
// kicked vhat.  R: records written on the way (a folded reset's).
template <bool START>
__global__ void k_permute(int ncap, int n, const int* __restrict__ perm, Store src, Store dst,
                          const int* __restrict__ gsrc, const int* __restrict__ gord,
                          const int* __restrict__ gcell, Box B, int* __restrict__ inv,
                          const int* __restrict__ ngp, const DevGlobals* __restrict__ g,
                          WrapFold F, PermRecords R) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && R.lists) R.lflag[(R.gen + 1) & 1] = 0;  // the next generation's flag
  if (p >= liveEntries(ncap, n, ngp)) return;
  const int e = perm[p];
  int i = e;
  const double dt = START ? g->dt / 2 : 0.0;
  if (e < n) {
    inv[e] = p;
  } else {
    i = gsrc[e - n];
    int sx, sy, sz;
    ordShift(gord[e - n], sx, sy, sz);
    const double4 q = src.posh[i];
    const Img im = imagePos(q, sx, sy, sz, B);
    const int t = src.iam[i];
    double4 vh = src.vhm[i];
    if (START && isSphT(t) && t != T_SPHB) {
      const double4 f = src.frc[i];
      halfKick(vh, f, dt);
    }
    const double4 qi = make_double4(im.x, im.y, im.z, q.w);
    dst.posh[p] = qi;
    dst.vhm[p] = vh;
    dst.dpo[p] = src.dpo[i];
    dst.iam[p] = t;
    dst.orig[p] = src.orig[i];
    dst.cell[p] = gcell[e - n];
    permRecords(R, p, qi, vh);
    return;  // the other fields of a ghost are never read
  }
  const int t = src.iam[i];
  const double4 q = src.posh[i];
  dst.posh[p] = q;
  dst.dpo[p] = src.dpo[i];
  if (p < R.ns) {
    double4 vh = src.vhm[i];
    if (START && isSphT(t) && t != T_SPHB) {
      const double4 f = src.frc[i];
      halfKick(vh, f, dt);
    }
    permRecords(R, p, q, vh);
  }
  if (START && isSphT(t)) {
    startKick<true>(src, i, dst, p, t, dt, F.bodyMask, F.ax, F.ay, F.az);
  } else {
    dst.vhm[p] = src.vhm[i];
    dst.vel[p] = src.vel[i];
    dst.v0p[p] = src.v0p[i];
    dst.vh0d[p] = src.vh0d[i];
    dst.frc[p] = src.frc[i];
    dst.fp[p] = src.fp[i];
    dst.fv[p] = src.fv[i];
    dst.fb[p] = src.fb[i];
  }
  dst.fdrag[p] = src.fdrag[i];
  dst.fvel[p] = src.fvel[i];
  dst.iam[p] = t;
  dst.orig[p] = src.orig[i];
  dst.cell[p] = src.cell[i];
}

// For the ghost at store position n + g: its origin's new store position and
// its image code (sx+1) + 3(sy+1) + 9(sz+1).
__global__ void k_ghost_link(int n, int ng, const int* __restrict__ perm,
                             const int* __restrict__ inv, const int* __restrict__ gsrc,
                             const int* __restrict__ gord, int* __restrict__ gorig,
                             int* __restrict__ gcode, const int* __restrict__ ngp) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= liveEntries(ng, 0, ngp)) return;
  const int e = perm[n + g];
  gorig[g] = inv[gsrc[e - n]];
  int sx, sy, sz;
  ordShift(gord[e - n], sx, sy, sz);
  gcode[g] = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
}
// Ghost fields the pair passes read (position and h, vhat and mass, dens,
// pdr2, omega, porosity) refreshed from their origins: the values a
// syncParticlesBetweenProcs would ship.
__global__ void k_ghost_refresh(Store S, int n, int ng, const int* __restrict__ gorig,
                                const int* __restrict__ gcode, Box B,
                                const int* __restrict__ ngp) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= liveEntries(ng, 0, ngp)) return;
  const int o = gorig[g];
  const int code = gcode[g];
  const double4 q = S.posh[o];
  const Img im = imagePos(q, code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1, B);
  S.posh[n + g] = make_double4(im.x, im.y, im.z, q.w);
  S.vhm[n + g] = S.vhm[o];
  S.dpo[n + g] = S.dpo[o];
}

// ---- a reset without host round trips (steps replayed from a graph) -------------
// gate[0]: set once a replayed step's reset finds what only the host can
// handle (its grid would change: hmax outside [0.9 ssr, ssr]; particles were
// deleted; more ghosts than the step processes; an error flag); it stays set,
// and gate[1] keeps gate[2], the replayed step's number, at that moment.
__global__ void k_reset_gate(const DevGlobals* __restrict__ g, int* __restrict__ gate, double ssr,
                             int keepSsr, unsigned long long ndelSph, unsigned long long ndelDem,
                             int ghostCap) {
  if (gate[0]) return;
  const double hmax = __longlong_as_double((long long)g->hmax_bits);
  bool stop = g->errflags != 0u || g->ndel_sph != ndelSph || g->ndel_dem != ndelDem ||
              g->ghostTotal > ghostCap;
  if (!keepSsr && (hmax > ssr || hmax < 0.9 * ssr)) stop = true;
  if (stop) {
    gate[0] = 1;
    gate[1] = gate[2];
  }
}
// The state a replayed step starts from (every stream of the real entries
// and the device globals), kept until a later replay starts cleanly; the
// first thread also counts the replays.  Nothing is copied once the gate is
// set, so the copy holds the state before the step that set it.
__global__ void k_snapshot(Store src, Store dst, int n, const DevGlobals* __restrict__ g,
                           DevGlobals* __restrict__ gkeep, int* __restrict__ gate) {
  if (gate[0]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *gkeep = *g;
    gate[2]++;
  }
  if (i >= n) return;
  dst.posh[i] = src.posh[i];
  dst.vhm[i] = src.vhm[i];
  dst.dpo[i] = src.dpo[i];
  dst.vel[i] = src.vel[i];
  dst.v0p[i] = src.v0p[i];
  dst.vh0d[i] = src.vh0d[i];
  dst.frc[i] = src.frc[i];
  dst.fp[i] = src.fp[i];
  dst.fv[i] = src.fv[i];
  dst.fb[i] = src.fb[i];
  dst.fdrag[i] = src.fdrag[i];
  dst.fvel[i] = src.fvel[i];
  dst.iam[i] = src.iam[i];
  dst.orig[i] = src.orig[i];
  dst.cell[i] = src.cell[i];
}

// ============================================================ SPH pair passes
// The neighbour predicate is Cparticle::is_neighbr (particle.h:236-238):
// b != a && len2(r_a - r_b) <= pow(2*max(h_a,h_b), 2); hq holds the h the
// cached lists were built with.

// CsphIncompress::calcPressSpsoundPdr2 sphIncompress.h:863-896
// D (optional): also the mixed form's D record {dens, pdr2, 1/h, 1/h^5} of
// each entry, as the mixed record packing would write it afterwards.
__global__ void k_eos(Store S, int ns, sphdem_params P, float4* __restrict__ D) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns) return;
  if (!isSphT(S.iam[a])) return;
  double4 d = S.dpo[a];
  double press;
  if (P.flags & SPHDEM_F_LIQ_DEM)
    press = P.prb * (pow(d.x / (P.refd * d.w), P.gamma) - 1.0);
  else
    press = P.prb * (pow(d.x / P.refd, 7.0) - 1.0);
  // only pdr2 and press are stored (omega may be written beside this pass)
  const double pdr2 = press / (d.x * d.x);
  S.dpo[a].y = pdr2;
  S.v0p[a].w = press;
  if (D) {
    const double rh = 1.0 / S.posh[a].w;
    D[a] = make_float4((float)d.x, (float)pdr2, (float)rh, (float)(rh * rh * rh * rh * rh));
  }
}

// Nsph::accelCondition sph.h:566-568 over sph, sphBoundary and dem
// dt rule (sphIncompress.cpp:126-154): min with LIQ_DEM_TIMESTEP and, after
// TIME_DROP_PARTICLE, DEM_TIMESTEP; both branches set dt = newDt.
__global__ void k_dt_final(sphdem_params P, DevGlobals* g) {
  double nd = __longlong_as_double((long long)g->newDt_bits);
  if (P.flags & SPHDEM_F_LIQ_DEM) {
    nd = dmin2(nd, P.liq_dem_timestep);
    if (!(P.flags & (SPHDEM_F_FIXED_DEM | SPHDEM_F_NO_DEM_CONTACTS)))
      if (g->time > P.time_drop) nd = dmin2(nd, P.dem_timestep);
  }
  g->newDt = nd;
  g->oldDt = g->dt;
  g->dt = nd;
}

// accelCondition over sph, sphBoundary and dem
// (grid-stride over a few hundred blocks: few atomics on newDt)
__global__ void k_accel(Store S, int n, DevGlobals* g) {
  double v = __longlong_as_double(0x7ff0000000000000ll);
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
    const int t = S.iam[a];
    if (isSphT(t) || t == T_DEM) {
      const double4 f = S.frc[a];
      v = dmin2(v, 0.25 * sqrt(S.posh[a].w / sqrt(dot3(f.x, f.y, f.z, f.x, f.y, f.z))));
    }
  }
  atomicMinPos(&g->newDt_bits, v);
}

// Nsph::calcHFromDens sph.h:1823-1825
__global__ void k_hfromdens(Store S, int ns, double hfac) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns || !isSphT(S.iam[a])) return;
  const double h = hfac * pow_third(S.vhm[a].w / S.dpo[a].x);
  S.posh[a].w = h;
}

// The candidate ranges of forCells for a wave-uniform query point: visit(j0,
// j1, sx, sy, sz) once per column range, in forCells' order.
template <class V>
__device__ __forceinline__ void forCellRanges(const Grid& g, const int* __restrict__ start,
                                              const Box& B, double qx, double qy, double qz,
                                              int reach, V&& visit) {
  const int nsx = B.periodic[0] ? 3 : 1, nsy = B.periodic[1] ? 3 : 1, nsz = B.periodic[2] ? 3 : 1;
  const double q[3] = {qx, qy, qz};
  for (int ix = 0; ix < nsx; ix++) {
    const int sx = ix == 0 ? 0 : (ix == 1 ? -1 : 1);
    const double tx = sx == 0 ? q[0] : (sx == 1 ? q[0] - B.Lp[0] : q[0] - B.Lm[0]);
    const int cx = cellOf(tx, g, 0);
    const int x0 = max(cx - reach, 0), x1 = min(cx + reach, g.dim[0] - 1);
    if (x0 > x1) continue;
    for (int iy = 0; iy < nsy; iy++) {
      const int sy = iy == 0 ? 0 : (iy == 1 ? -1 : 1);
      const double ty = sy == 0 ? q[1] : (sy == 1 ? q[1] - B.Lp[1] : q[1] - B.Lm[1]);
      const int cy = cellOf(ty, g, 1);
      const int y0 = max(cy - reach, 0), y1 = min(cy + reach, g.dim[1] - 1);
      if (y0 > y1) continue;
      for (int iz = 0; iz < nsz; iz++) {
        const int sz = iz == 0 ? 0 : (iz == 1 ? -1 : 1);
        const double tz = sz == 0 ? q[2] : (sz == 1 ? q[2] - B.Lp[2] : q[2] - B.Lm[2]);
        const int cz = cellOf(tz, g, 2);
        const int z0 = max(cz - reach, 0), z1 = min(cz + reach, g.dim[2] - 1);
        if (z0 > z1) continue;
        for (int a = x0; a <= x1; a++)
          for (int b = y0; b <= y1; b++) {
            const int k0 = g.base + (a * g.dim[1] + b) * g.dim[2];
            visit(start[k0 + z0], start[k0 + z1 + 1], sx, sy, sz);
          }
      }
    }
  }
}

// Accepted candidates of one wavefront, compacted in LDS so that the pair
// arithmetic runs on full wavefronts: lanes test 64 consecutive candidates,
// the accepted ones are appended in lane order (ballot + prefix count), and
// every 64 appended entries are processed together, lane i taking entry i.
// The order of the entries, and so of every lane's partial sums, depends
// only on the candidate order: results are the same run to run.
constexpr int kCompactCap = 128;  // 64 carried over + 64 appended
// ranges(visit) calls visit(j0, j1, code) for each candidate range (code: an
// image code the body receives with the candidate); test(j, code) decides a
// candidate, body(j, code) runs for each accepted one.
template <class Ranges, class Test, class Body>
__device__ __forceinline__ void compactedPairs(int* __restrict__ lb, int* __restrict__ lc,
                                               int lane, Ranges&& ranges, Test&& test,
                                               Body&& body) {
  int cnt = 0;  // wave-uniform
  auto flush = [&](int m) {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < m) body(lb[lane], lc[lane]);
    __builtin_amdgcn_wave_barrier();
  };
  ranges([&](int j0, int j1, int code) {
    for (int jj = j0; jj < j1; jj += 64) {
      const int j = jj + lane;
      const bool ok = j < j1 && test(j, code);
      const unsigned long long m = __ballot(ok);
      if (ok) {
        const int at = cnt + (int)__builtin_amdgcn_mbcnt_hi(
                                 (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        lb[at] = j;
        lc[at] = code;
      }
      cnt += __popcll(m);
      if (cnt >= 64) {
        flush(64);
        // carry the entries past the first 64 to the front
        int cb = 0, cc = 0;
        if (lane < cnt - 64) {
          cb = lb[64 + lane];
          cc = lc[64 + lane];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < cnt - 64) {
          lb[lane] = cb;
          lc[lane] = cc;
        }
        cnt -= 64;
      }
    }
  });
  if (cnt > 0) flush(cnt);
}
__device__ __forceinline__ int imageCode(int sx, int sy, int sz) {
  return (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
}
__device__ __forceinline__ void imageShift(int code, int& sx, int& sy, int& sz) {
  sx = code % 3 - 1;
  sy = (code / 3) % 3 - 1;
  sz = code / 9 - 1;
}

// ============================================================ SPH <-> DEM coupling
// calcNeighboursCoupling's SPH-side rule (dataLL.cpp:1154-1190) as a predicate
// on (DEM cell of the SPH particle, DEM cell of the DEM image, distance).
struct CouplingRule {
  double cmin[3];
  double e;       // 2*DEM_SEARCH_RADIUS
  double einv;    // 1/e (0: every cell index takes the division)
  double s;       // 2*LIQ_DEM_COUPLING_RADIUS/DEM_SEARCH_RADIUS
  double thr;     // pow(0.5*Rc/DSR, 2)
  double rc2;     // pow(Rc, 2)
  double skip2;   // pairs beyond this squared distance have W == 0 exactly
  int nc;         // ceil(0.5*Rc/DSR)
  // DEM ghost band (dataLL.cpp:337-364): images exist only within Rc of a face
  double gmin[3], gspace[3];
};
// the fast product as cellTrunc; the division, where needed, in the
// reference's own order ((x - cmin)/e + s) - 1, which rounds differently from
// (x - cmin)/e + (s - 1) exactly at the cell boundaries where it is taken
__device__ __forceinline__ int demCell(const CouplingRule& R, double x, int d) {
  const double t = (x - R.cmin[d]) * R.einv + (R.s - 1);
  if (fabs(t - rint(t)) > 1e-9) return (int)t;
  return (int)((x - R.cmin[d]) / R.e + R.s - 1);
}
// accepted(l2) with the whole-cell (addAll) branch; `full` keeps addAll pairs
// whatever their distance (the reference's list), otherwise only pairs that
// can carry a nonzero kernel value are reported.
__device__ __forceinline__ bool couplingAccept(const CouplingRule& R, const int* ca, double ix,
                                               double iy, double iz, double l2, bool full) {
  const int d0 = demCell(R, ix, 0) - ca[0];
  const int d1 = demCell(R, iy, 1) - ca[1];
  const int d2 = demCell(R, iz, 2) - ca[2];
  if (d0 < -R.nc || d0 > R.nc || d1 < -R.nc || d1 > R.nc || d2 < -R.nc || d2 > R.nc) return false;
  const int plus = (d0 + 1) * (d0 + 1) + (d1 + 1) * (d1 + 1) + (d2 + 1) * (d2 + 1);
  const int minus = (d0 - 1) * (d0 - 1) + (d1 - 1) * (d1 - 1) + (d2 - 1) * (d2 - 1);
  if (plus < R.thr) return full || l2 <= R.skip2;
  if (minus < R.thr) return l2 < R.rc2;
  return false;
}
__device__ __forceinline__ bool demImageExists(const CouplingRule& R, const Box& B,
                                               const double4& p, int sx, int sy, int sz) {
  const int s[3] = {sx, sy, sz};
  for (int d = 0; d < 3; d++) {
    if (s[d] == 0) continue;
    if (!B.periodic[d]) return false;
    int ic = (int)((comp(const_cast<double4&>(p), d) - R.gmin[d]) / R.gspace[d] + 1);
    ic = ic > 2 ? 2 : (ic < 0 ? 0 : ic);
    if (ic != (s[d] == 1 ? 0 : 2)) return false;
  }
  return true;
}

// Shepard sum of each DEM particle (calcShepSum, sph.h:1303-1320, a scatter
// from SPH in the reference, followed by reverseSync) as a gather over the SPH
// particles whose coupling list holds this DEM particle or one of its images.
// One wavefront per DEM particle; the lanes split the candidates and the
// partial sums meet in a fixed butterfly.
__global__ void __launch_bounds__(256) k_shepsum(Store S, int ns, int nd, Grid gs,
                                                 const int* __restrict__ start, Box B,
                                                 CouplingRule R, int reach, double wcon,
                                                 double Rc) {
  __shared__ int lbuf[4][kCompactCap], lcode[4][kCompactCap];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = ns + xcdBlock() * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= ns + nd) return;  // uniform per wavefront
  const double4 pj = S.posh[j];
  double sum = 0.0;
  Box B0 = B;
  B0.periodic[0] = B0.periodic[1] = B0.periodic[2] = 0;
  // candidates a see the image r_j + s L; enumerate them around that image
  auto ranges = [&](auto&& visit) {
    const int nsx = B.periodic[0] ? 3 : 1, nsy = B.periodic[1] ? 3 : 1,
              nsz = B.periodic[2] ? 3 : 1;
    for (int ix = 0; ix < nsx; ix++)
      for (int iy = 0; iy < nsy; iy++)
        for (int iz = 0; iz < nsz; iz++) {
          const int sx = ix == 0 ? 0 : (ix == 1 ? -1 : 1);
          const int sy = iy == 0 ? 0 : (iy == 1 ? -1 : 1);
          const int sz = iz == 0 ? 0 : (iz == 1 ? -1 : 1);
          if ((sx | sy | sz) != 0 && !demImageExists(R, B, pj, sx, sy, sz)) continue;
          const Img im = imagePos(pj, sx, sy, sz, B);
          forCellRanges(gs, start, B0, im.x, im.y, im.z, reach,
                        [&](int j0, int j1, int, int, int) { visit(j0, j1, imageCode(sx, sy, sz)); });
        }
  };
  auto dist2 = [&](int a, int code, Img& im) {
    int sx, sy, sz;
    imageShift(code, sx, sy, sz);
    im = imagePos(pj, sx, sy, sz, B);
    const double4 pa = S.posh[a];
    const double dx = pa.x - im.x, dy = pa.y - im.y, dz = pa.z - im.z;
    return dot3(dx, dy, dz, dx, dy, dz);
  };
  // the test keeps the cheap distance cut; the coupling cell rule runs on the
  // compacted candidates
  compactedPairs(lbuf[w], lcode[w], lane, ranges, [&](int a, int code) {
    if (!isSphT(S.iam[a])) return false;
    Img im;
    return dist2(a, code, im) <= R.skip2;
  }, [&](int a, int code) {
    Img im;
    const double l2 = dist2(a, code, im);
    const double4 pa = S.posh[a];
    const int ca[3] = {demCell(R, pa.x, 0), demCell(R, pa.y, 1), demCell(R, pa.z, 2)};
    if (!couplingAccept(R, ca, im.x, im.y, im.z, l2, false)) return;
    const double dv = S.vhm[a].w / S.dpo[a].x;
    const double r = sqrt(l2);
    sum += dv * wendW(wcon, r * (2.0 / Rc), Rc / 2.0);
  });
  sum = waveSum(sum);
  if (lane == 0) S.fp[j].w = sum;
}

// SPH side of middle()'s coupling block (sphIncompress.cpp:64-78): calcPorosity
// (sph.h:1323-1341), calcDEMtoSPHDrag (sph.h:1343-1362) and finalisePorosity
// (sph.h:1289-1296), one gather over the DEM partners.
// DEM cells with a DEM particle within one cell (the 3x3x3 block around them):
// near[c] == gen.  A coupling search centred on any other cell of the DEM grid
// finds no candidate and is skipped.
__global__ void k_bump(int* __restrict__ p) { ++*p; }
__global__ void k_dem_near(Store S, int ns, int nd, Grid gc, int* __restrict__ near,
                           const int* __restrict__ genp) {
  const int j = ns + blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns + nd || S.iam[j] != T_DEM) return;
  const int gen = *genp;
  const int c = S.cell[j];
  const int x = c & 2047, y = (c >> 11) & 2047, z = cellZ(c);
  for (int a = max(x - 1, 0); a <= min(x + 1, gc.dim[0] - 1); a++)
    for (int b = max(y - 1, 0); b <= min(y + 1, gc.dim[1] - 1); b++)
      for (int d = max(z - 1, 0); d <= min(z + 1, gc.dim[2] - 1); d++)
        near[(a * gc.dim[1] + b) * gc.dim[2] + d] = gen;
}

__global__ void k_coupling_sph(Store S, int ns, int nd, Grid gc, const int* __restrict__ start,
                               Box B, CouplingRule R, sphdem_params P,
                               const int* __restrict__ near, const int* __restrict__ genp) {
  const int a = xcdThread();
  if (a >= ns || !isSphT(S.iam[a])) return;
  const int gen = *genp;
  const double4 pa = S.posh[a];
  const double Rc = P.coupling_radius;
  const double rdens = 1.0 / S.dpo[a].x;
  const int ca[3] = {demCell(R, pa.x, 0), demCell(R, pa.y, 1), demCell(R, pa.z, 2)};
  double por = 0.0;
  double fd0 = 0.0, fd1 = 0.0, fd2 = 0.0;
  // forCells over the DEM grid, skipping each image shift whose centre cell
  // lies in the grid with no DEM particle around it
  auto keep = [&](int cx, int cy, int cz) {
    if (cx < 0 || cy < 0 || cz < 0 || cx >= gc.dim[0] || cy >= gc.dim[1] || cz >= gc.dim[2])
      return true;
    return near[(cx * gc.dim[1] + cy) * gc.dim[2] + cz] == gen;
  };
  forCells(gc, start, B, pa.x, pa.y, pa.z, 1, [&](int j, int sx, int sy, int sz) {
    const double4 pj = S.posh[j];
    if ((sx | sy | sz) != 0 && !demImageExists(R, B, pj, sx, sy, sz)) return;
    const Img im = imagePos(pj, sx, sy, sz, B);
    const double dx = pa.x - im.x, dy = pa.y - im.y, dz = pa.z - im.z;
    const double l2 = dot3(dx, dy, dz, dx, dy, dz);
    if (l2 > R.skip2) return;
    if (!couplingAccept(R, ca, im.x, im.y, im.z, l2, false)) return;
    const double r = sqrt(l2);
    const double w = wendW(P.wcon, r * (2.0 / Rc), Rc / 2.0);
    por += P.porosity_incr * w;
    const double4 fdj = S.fdrag[j];
    const double s = S.vhm[j].w * rdens;
    const double shep = S.fp[j].w;
    fd0 -= s * fdj.x * w / shep;
    fd1 -= s * fdj.y * w / shep;
    fd2 -= s * fdj.z * w / shep;
  }, 0, 1, keep);
  S.fdrag[a] = make_double4(fd0, fd1, fd2, 0.0);
  double4 f = S.frc[a];
  f.x += fd0;
  f.y += fd1;
  f.z += fd2;
  S.frc[a] = f;
  por = 1.0 - por;
  if (por < 0.37) por = 0.37;
  S.dpo[a].w = por;
}

// calcFPI (sph.h:1708-1792) with calcParticleDrag (sph.h:1575-1631, Di Felice):
// each DEM particle gathers its SPH partners (the 27 SPH cells around it,
// dataLL.cpp:1106-1126).
// One wavefront per DEM particle: the lanes test the candidates, the accepted
// ones are compacted and summed on full wavefronts, and the partial sums meet
// in a fixed butterfly.
__global__ void __launch_bounds__(256) k_fpi(Store S, int ns, int nd, Grid gs,
                                             const int* __restrict__ start, Box B,
                                             sphdem_params P) {
  __shared__ int lbuf[4][kCompactCap], lcode[4][kCompactCap];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = ns + xcdBlock() * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= ns + nd) return;  // uniform per wavefront
  const double4 pj = S.posh[j];
  const double4 vj = S.vhm[j];
  double sum = 0, fdens = 0, por = 0;
  double ff[3] = {0, 0, 0}, fpfb[3] = {0, 0, 0}, fv[3] = {0, 0, 0}, fdr[3] = {0, 0, 0},
         fvel[3] = {0, 0, 0};
  auto dist2 = [&](const double4& pb, int sx, int sy, int sz) {
    const Img im = imagePos(pb, sx, sy, sz, B);
    const double dx = pj.x - im.x, dy = pj.y - im.y, dz = pj.z - im.z;
    return dot3(dx, dy, dz, dx, dy, dz);
  };
  auto ranges = [&](auto&& visit) {
    forCellRanges(gs, start, B, pj.x, pj.y, pj.z, 1, [&](int j0, int j1, int sx, int sy, int sz) {
      visit(j0, j1, imageCode(sx, sy, sz));
    });
  };
  compactedPairs(lbuf[w], lcode[w], lane, ranges, [&](int b, int code) {
    int sx, sy, sz;
    imageShift(code, sx, sy, sz);
    const double4 pb = S.posh[b];
    const double lim = 2.0 * dmax2(pj.w, pb.w);
    return dist2(pb, sx, sy, sz) <= lim * lim && S.iam[b] != T_BOUNDARY;
  }, [&](int b, int code) {
    int sx, sy, sz;
    imageShift(code, sx, sy, sz);
    const double4 pb = S.posh[b];
    const double r = sqrt(dist2(pb, sx, sy, sz));
    const double4 vb = S.vhm[b];
    // dens and porosity only: omega may be written meanwhile (the fused dddt walk)
    const double2 db = make_double2(S.dpo[b].x, S.dpo[b].w);
    const double dvol = vb.w / db.x;
    const double q = r / pb.w;
    const double Wab = wendW(P.wcon, q, pb.w);
    const double dvWab = dvol * Wab;
    const double mW = vb.w * Wab;
    const double4 fpb = S.fp[b], fbb = S.fb[b], fvb = S.fv[b], fb_ = S.frc[b], fdb = S.fdrag[b];
    fpfb[0] += mW * (fpb.x + fbb.x);
    fpfb[1] += mW * (fpb.y + fbb.y);
    fpfb[2] += mW * (fpb.z + fbb.z);
    fv[0] += mW * (fvb.x);
    fv[1] += mW * (fvb.y);
    fv[2] += mW * (fvb.z);
    fv[0] += dvWab * (fvb.x);
    fv[1] += dvWab * (fvb.y);
    fv[2] += dvWab * (fvb.z);
    ff[0] += dvWab * (fb_.x);
    ff[1] += dvWab * (fb_.y);
    ff[2] += dvWab * (fb_.z);
    fdr[0] += dvWab * fdb.x;
    fdr[1] += dvWab * fdb.y;
    fdr[2] += dvWab * fdb.z;
    fvel[0] += dvWab * vb.x;
    fvel[1] += dvWab * vb.y;
    fvel[2] += dvWab * vb.z;
    por += dvWab * db.y;
    fdens += Wab * vb.w;
    sum += dvWab;
  });
  sum = waveSum(sum);
  fdens = waveSum(fdens);
  por = waveSum(por);
  for (int k = 0; k < 3; k++) {
    ff[k] = waveSum(ff[k]);
    fpfb[k] = waveSum(fpfb[k]);
    fv[k] = waveSum(fv[k]);
    fdr[k] = waveSum(fdr[k]);
    fvel[k] = waveSum(fvel[k]);
  }
  if (lane != 0) return;
  double4 fdrag_out = make_double4(0.0, 0.0, 0.0, 0.0);
  double4 f = S.frc[j];
  if (sum > 0.5) {
    const double rsum = 1.0 / sum;
    for (int k = 0; k < 3; k++) {
      ff[k] *= rsum;
      fpfb[k] *= rsum;
      fv[k] *= rsum;
      fdr[k] *= rsum;
      fvel[k] *= rsum;
    }
    fdens *= rsum;
    por *= rsum;
    S.fb[j] = make_double4(fpfb[0], fpfb[1], fpfb[2], 0.0);
    S.fv[j] = make_double4(fv[0], fv[1], fv[2], 0.0);
    // calcParticleDrag
    const double dv0 = por * (fvel[0] - vj.x), dv1 = por * (fvel[1] - vj.y),
                 dv2 = por * (fvel[2] - vj.z);
    const double vdiff = sqrt(dot3(dv0, dv1, dv2, dv0, dv1, dv2));
    double fd0 = 0.0, fd1 = 0.0, fd2 = 0.0;
    if (vdiff != 0) {
      const double A = P.pi * (P.dem_radius * P.dem_radius);
      const double Re = (2.0 * P.dem_radius / P.viscosity) * vdiff;
      const double t = 1.5 - log10(Re);
      const double dragGamma = P.drag_a - P.drag_b * exp(-(t * t) / 2.0);
      const double c = 0.63 + 4.8 / sqrt(Re);
      const double C = c * c;
      const double s = (1.0 / 2.0) * A * C * vdiff;
      const double pw = pow(por, -dragGamma);
      fd0 = s * dv0 * pw;
      fd1 = s * dv1 * pw;
      fd2 = s * dv2 * pw;
    }
    const double dm = P.dens / vj.w;
    const double sv = P.dem_vol / vj.w;
    fdrag_out.x = sv * (fv[0] + fpfb[0]) + dm * fd0;
    fdrag_out.y = sv * (fv[1] + fpfb[1]) + dm * fd1;
    fdrag_out.z = sv * (fv[2] + fpfb[2]) + dm * fd2;
    f.x += fdrag_out.x;
    f.y += fdrag_out.y;
    f.z += fdrag_out.z;
    S.frc[j] = f;
  }
  S.fdrag[j] = fdrag_out;
  S.fvel[j] = make_double4(fvel[0], fvel[1], fvel[2], 0.0);
  S.dpo[j].w = por;
  S.fp[j].w = sum;
}

// calcDEMContacts (sph.h:1676-1706) with demNormal LINEAR (sph.h:1513-1515)
// One thread per sphere walking its cells in order (the reference's loop as
// written): with tens of thousands of spheres it shares the GPU with the
// passes beside it better than a wavefront per sphere does.
__global__ void k_contacts_lane(Store S, int ns, int nd, Grid gc, const int* __restrict__ start,
                           Box B, sphdem_params P) {
  const int j = ns + blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns + nd) return;
  const double4 pj = S.posh[j];
  const double4 vj = S.vhm[j];
  double4 fp = S.fp[j], f = S.frc[j];
  bool any = false;
  forCells(gc, start, B, pj.x, pj.y, pj.z, 1, [&](int b, int sx, int sy, int sz) {
    if (b == j && (sx | sy | sz) == 0) return;
    const double4 pb = S.posh[b];
    const Img im = imagePos(pb, sx, sy, sz, B);
    const double dx = pj.x - im.x, dy = pj.y - im.y, dz = pj.z - im.z;
    const double l2 = dot3(dx, dy, dz, dx, dy, dz);
    const double lim = 2.0 * dmax2(pj.w, pb.w);
    if (!(l2 <= lim * lim)) return;
    const double4 vb = S.vhm[b];
    const double r = sqrt(l2);
    const double dv0 = vj.x - vb.x, dv1 = vj.y - vb.y, dv2 = vj.z - vb.z;
    const double n0 = dx / r, n1 = dy / r, n2 = dz / r;
    const double overlap = 2.0 * P.dem_radius - r;
    if (overlap > 0) {
      const double od = -dot3(dv0, dv1, dv2, n0, n1, n2);
      const double s = P.dem_k * overlap + P.dem_gamma * od;
      const double f0 = s * n0 / vj.w, f1 = s * n1 / vj.w, f2 = s * n2 / vj.w;
      fp.x += f0;
      fp.y += f1;
      fp.z += f2;
      f.x += f0;
      f.y += f1;
      f.z += f2;
      any = true;
    }
  });
  if (any) {
    S.fp[j] = fp;
    S.frc[j] = f;
  }
}

__global__ void __launch_bounds__(256) k_contacts(Store S, int ns, int nd, Grid gc,
                                                  const int* __restrict__ start, Box B,
                                                  sphdem_params P) {
  // one wavefront per sphere: the lanes test the candidates of its cells 64 at
  // a time (the cells' chain of dependent loads then runs once per column
  // range, not once per candidate), the partners within reach are appended in
  // candidate order, and lane 0 applies their contact forces one after
  // another in that order -- the sums of the one-thread walk, bit for bit
  __shared__ int lbuf[4][kCompactCap], lcode[4][kCompactCap];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = ns + xcdBlock() * (blockDim.x >> 6) + w;
  if (j >= ns + nd) return;  // uniform per wavefront
  const double4 pj = S.posh[j];
  const double4 vj = S.vhm[j];
  double4 fp = S.fp[j], f = S.frc[j];
  bool any = false;
  auto ranges = [&](auto&& visit) {
    forCellRanges(gc, start, B, pj.x, pj.y, pj.z, 1, [&](int j0, int j1, int sx, int sy, int sz) {
      visit(j0, j1, imageCode(sx, sy, sz));
    });
  };
  auto dist2 = [&](const double4& pb, int code) {
    int sx, sy, sz;
    imageShift(code, sx, sy, sz);
    const Img im = imagePos(pb, sx, sy, sz, B);
    const double dx = pj.x - im.x, dy = pj.y - im.y, dz = pj.z - im.z;
    return make_double4(dx, dy, dz, dot3(dx, dy, dz, dx, dy, dz));
  };
  int cnt = 0;  // wave-uniform
  auto flush = [&](int m) {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0)
      for (int k = 0; k < m; k++) {
        const int b = lbuf[w][k];
        const double4 d = dist2(S.posh[b], lcode[w][k]);
        const double4 vb = S.vhm[b];
        const double r = sqrt(d.w);
        const double dv0 = vj.x - vb.x, dv1 = vj.y - vb.y, dv2 = vj.z - vb.z;
        const double n0 = d.x / r, n1 = d.y / r, n2 = d.z / r;
        const double overlap = 2.0 * P.dem_radius - r;
        if (overlap > 0) {
          const double od = -dot3(dv0, dv1, dv2, n0, n1, n2);
          const double s = P.dem_k * overlap + P.dem_gamma * od;
          const double f0 = s * n0 / vj.w, f1 = s * n1 / vj.w, f2 = s * n2 / vj.w;
          fp.x += f0;
          fp.y += f1;
          fp.z += f2;
          f.x += f0;
          f.y += f1;
          f.z += f2;
          any = true;
        }
      }
    __builtin_amdgcn_wave_barrier();
  };
  ranges([&](int j0, int j1, int code) {
    for (int jj = j0; jj < j1; jj += 64) {
      const int b = jj + lane;
      bool ok = false;
      if (b < j1 && !(b == j && code == 13)) {
        const double4 pb = S.posh[b];
        const double lim = 2.0 * dmax2(pj.w, pb.w);
        ok = dist2(pb, code).w <= lim * lim;
      }
      const unsigned long long m = __ballot(ok);
      if (ok) {
        const int at = cnt + (int)__builtin_amdgcn_mbcnt_hi(
                                 (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        lbuf[w][at] = b;
        lcode[w][at] = code;
      }
      cnt += __popcll(m);
      if (cnt >= 64) {
        flush(cnt);
        cnt = 0;
      }
    }
  });
  if (cnt > 0) flush(cnt);
  if (lane == 0 && any) {
    S.fp[j] = fp;
    S.frc[j] = f;
  }
}

// ============================================================ integrator
// start(): driftRAndKick + driftRest + initSumsMiddle on sph/sphBoundary,
// driftR on dem (sphIncompress.cpp:4-16,37; sphIncompress.h:620-704, 744-763);
// time += dt/2.
// bodyMask/ax..az: a body force added right after initSumsMiddle (the
// beforeMiddle hook of run.cpp:112-113, addGravity) for the types in bodyMask.
__global__ void k_start(Store S, int n, sphdem_params P, DevGlobals* g, unsigned bodyMask,
                        double ax, double ay, double az) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const double dt = g->dt / 2;
  if (a == 0) g->time += g->dt / 2;
  if (a >= n) return;
  const int t = S.iam[a];
  if (!isSphT(t) && t != T_DEM) return;
  double4 p = S.posh[a];
  startDrift(S, a, t, P.flags, P.hfac, dt, p);
  S.posh[a] = p;
  if (isSphT(t)) startKick<false>(S, a, S, a, t, dt, bodyMask, ax, ay, az);
}


// kick (sphIncompress.h:706-742) for sph with (oldDt/2, dt/2), initSumsEnd on
// sph/sphBoundary, then driftRAndKick + initSumsMiddle on dem
// (sphIncompress.cpp:157-163).
__global__ void k_mid_kick(Store S, int n, sphdem_params P, DevGlobals* g) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int t = S.iam[a];
  if (isSphT(t)) {
    if (t == T_SPH) {
      const double dt = 0.5 * g->oldDt + 0.5 * g->dt;
      const double4 f = S.frc[a];
      const double4 v0 = S.v0p[a];
      const double4 vh0 = S.vh0d[a];
      double4 v = S.vel[a];
      double4 vh = S.vhm[a];
      v.x = v0.x + dt * f.x;
      v.y = v0.y + dt * f.y;
      v.z = v0.z + dt * f.z;
      vh.x = vh0.x + dt * f.x;
      vh.y = vh0.y + dt * f.y;
      vh.z = vh0.z + dt * f.z;
      if (g->time < P.damptime) {
        v.x *= 0.98;
        v.y *= 0.98;
        v.z *= 0.98;
        vh.x *= 0.98;
        vh.y *= 0.98;
        vh.z *= 0.98;
      }
      S.vel[a] = v;
      S.vhm[a] = vh;
    }
    S.vh0d[a].w = 0.0;  // initSumsEnd
  } else if (t == T_DEM && (P.flags & SPHDEM_F_LIQ_DEM)) {
    const double dt = g->dt / 2;
    double4 p = S.posh[a];
    double4 vh = S.vhm[a];
    double4 v = S.vel[a];
    const double4 f = S.frc[a];
    p.x += dt * vh.x;
    p.y += dt * vh.y;
    p.z += dt * vh.z;
    S.v0p[a] = make_double4(v.x, v.y, v.z, S.v0p[a].w);
    S.vh0d[a] = make_double4(vh.x, vh.y, vh.z, S.vh0d[a].w);
    if (P.flags & SPHDEM_F_FIXED_DEM) {
      // v and vhat stay at v0, vhat0
    } else {
      v.x += dt * f.x;
      v.y += dt * f.y;
      v.z += dt * f.z;
      vh.x += dt * f.x;
      vh.y += dt * f.y;
      vh.z += dt * f.z;
    }
    S.posh[a] = p;
    S.vel[a] = v;
    S.vhm[a] = vh;
    const double4 z = make_double4(0.0, 0.0, 0.0, 0.0);
    S.frc[a] = make_double4(0.0, 0.0, 0.0, 0.0);  // dudt too
    S.fp[a] = make_double4(0.0, 0.0, 0.0, S.fp[a].w);
    S.fv[a] = z;
    S.fb[a] = z;
  }
}

// end(): driftR on sph/sphBoundary (sphIncompress.cpp:174), time += dt/2
__global__ void k_end_drift(Store S, int ns, DevGlobals* g) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const double dt = g->dt / 2;
  if (a == 0) g->time += g->dt / 2;
  if (a >= ns || !isSphT(S.iam[a])) return;
  double4 p = S.posh[a];
  const double4 vh = S.vhm[a];
  p.x += dt * vh.x;
  p.y += dt * vh.y;
  p.z += dt * vh.z;
  S.posh[a] = p;
}

// kick on dem with (dt/2, dt/2) (sphIncompress.cpp:239)
// DEM_BLOCK (sphIncompress.cpp:231-238 with calcAvForce/setAvForce,
// sphIncompress.h:107-114): every DEM particle's f becomes the mean f of the
// DEM particles.  One workgroup: each thread sums a strided share of the DEM
// range in store order, a fixed tree adds the shares, then every thread writes
// the mean to its share (deterministic run to run).
constexpr int kDemBlockThreads = 1024;
__global__ void __launch_bounds__(kDemBlockThreads) k_dem_block(Store S, int ns, int nd) {
  __shared__ double red[3][kDemBlockThreads];
  double sx = 0.0, sy = 0.0, sz = 0.0, cnt = 0.0;
  for (int a = ns + threadIdx.x; a < ns + nd; a += kDemBlockThreads)
    if (S.iam[a] == T_DEM) {
      const double4 f = S.frc[a];
      sx += f.x;
      sy += f.y;
      sz += f.z;
      cnt += 1.0;
    }
  __shared__ double cred[kDemBlockThreads];
  red[0][threadIdx.x] = sx;
  red[1][threadIdx.x] = sy;
  red[2][threadIdx.x] = sz;
  cred[threadIdx.x] = cnt;
  __syncthreads();
  for (int s = kDemBlockThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
      cred[threadIdx.x] += cred[threadIdx.x + s];
    }
    __syncthreads();
  }
  const double n = cred[0];
  if (n == 0.0) return;
  const double ax = red[0][0] / n, ay = red[1][0] / n, az = red[2][0] / n;
  for (int a = ns + threadIdx.x; a < ns + nd; a += kDemBlockThreads)
    if (S.iam[a] == T_DEM) {
      S.frc[a].x = ax;
      S.frc[a].y = ay;
      S.frc[a].z = az;
    }
}

__global__ void k_dem_kick(Store S, int ns, int nd, sphdem_params P, DevGlobals* g) {
  const int a = ns + blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns + nd || S.iam[a] != T_DEM) return;
  const double dt = 0.5 * g->dt + 0.5 * g->dt;
  const double4 f = S.frc[a];
  const double4 v0 = S.v0p[a];
  const double4 vh0 = S.vh0d[a];
  double4 v = S.vel[a];
  double4 vh = S.vhm[a];
  v.x = v0.x + dt * f.x;
  v.y = v0.y + dt * f.y;
  v.z = v0.z + dt * f.z;
  vh.x = vh0.x + dt * f.x;
  vh.y = vh0.y + dt * f.y;
  vh.z = vh0.z + dt * f.z;
  if (g->time < P.damptime) {
    v.x *= 0.98;
    v.y *= 0.98;
    v.z *= 0.98;
    vh.x *= 0.98;
    vh.y *= 0.98;
    vh.z *= 0.98;
  }
  if (P.flags & SPHDEM_F_FIXED_DEM) {
    v.x = v0.x;
    v.y = v0.y;
    v.z = v0.z;
    vh.x = vh0.x;
    vh.y = vh0.y;
    vh.z = vh0.z;
  }
  S.vel[a] = v;
  S.vhm[a] = vh;
}

// driftRest + calcEnergies on sph/sphBoundary (sphIncompress.cpp:245-249;
// sphIncompress.h:678-704, 769-781)
__global__ void k_end_rest(Store S, int ns, sphdem_params P, DevGlobals* g) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns || !isSphT(S.iam[a])) return;
  const double dt = g->dt / 2;
  double4 d = S.dpo[a];
  const double dddt = S.vh0d[a].w;
  d.x += dt * dddt;
  if (!(P.flags & SPHDEM_F_CONST_H)) S.posh[a].w = P.hfac * pow_third(S.vhm[a].w / d.x);
  double press;
  if (P.flags & SPHDEM_F_LIQ_DEM)
    press = P.prb * (pow(d.x / (P.refd * d.w), P.gamma) - 1.0);
  else
    press = P.prb * (pow(d.x / P.refd, 7.0) - 1.0);
  d.y = press / (d.x * d.x);
  S.dpo[a] = d;
  S.v0p[a].w = press;
  const double dudt = d.y * dddt;
  S.frc[a].w = dudt;
  S.vel[a].w += g->dt * dudt;
  if (S.iam[a] == T_SPH) {  // the running viscous energy of calcEnergies
    const double4 vh = S.vhm[a], fv = S.fv[a];
    S.fvel[a].w += g->dt * -dot3(vh.x, vh.y, vh.z, fv.x, fv.y, fv.z);
  }
}

// ============================================================ hooks
__global__ void k_body_force(Store S, int n, unsigned mask, double ax, double ay, double az) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int t = S.iam[a];
  if (t >= 32 || !((mask >> t) & 1u)) return;
  double4 f = S.frc[a];
  if (ax != 0.0) f.x += ax;
  if (ay != 0.0) f.y += ay;
  if (az != 0.0) f.z += az;
  S.frc[a] = f;
}
// addCustomBoundaries sediment3D/customSim.cpp:67-81
__global__ void k_dem_floor(Store S, int ns, int nd, sphdem_params P) {
  const int a = ns + blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns + nd || S.iam[a] != T_DEM) return;
  const double4 p = S.posh[a];
  const double overlap = P.rmin[2] + P.dem_radius - p.z;
  if (overlap > 0.0) {
    const double od = -S.vhm[a].z;
    const double s = P.dem_k * overlap + P.dem_gamma * od;
    const double fz = s * 1.0 / S.vhm[a].w;
    S.fb[a].z += fz;
    S.frc[a].z += fz;
  }
}
// freezeParticles sediment3D/customSim.cpp:12-25, applied while time <= tfreeze
__global__ void k_freeze(Store S, int n, unsigned mask, double tfreeze, const DevGlobals* g) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n || !(g->time <= tfreeze)) return;
  const int t = S.iam[a];
  if (t >= 32 || !((mask >> t) & 1u)) return;
  S.frc[a] = make_double4(0.0, 0.0, 0.0, S.frc[a].w);
  S.vel[a] = make_double4(0.0, 0.0, 0.0, S.vel[a].w);
  S.vhm[a] = make_double4(0.0, 0.0, 0.0, S.vhm[a].w);
}
// sumDragForces sediment3D/customSim.cpp:27-42: per-block partial sums of
// m*fdrag for sph (0..2) and dem (3..5); k_sum_drag_final adds them in a fixed order.
__global__ void k_sum_drag(Store S, int n, double* partial) {
  __shared__ double red[6][256];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (a < n) {
    const int t = S.iam[a];
    const double m = S.vhm[a].w;
    const double4 fd = S.fdrag[a];
    if (t == T_SPH) {
      v[0] = m * fd.x;
      v[1] = m * fd.y;
      v[2] = m * fd.z;
    } else if (t == T_DEM) {
      v[3] = m * fd.x;
      v[4] = m * fd.y;
      v[5] = m * fd.z;
    }
  }
  for (int k = 0; k < 6; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) partial[6 * blockIdx.x + k] = red[k][0];
}
// one 256-thread block: strided partial sums in a fixed order, then a tree
__global__ void __launch_bounds__(256) k_sum_drag_final(const double* partial, int nb,
                                                        DevGlobals* g) {
  __shared__ double red[6][256];
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int k = 0; k < 6; k++) s[k] += partial[6 * b + k];
  for (int k = 0; k < 6; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) g->custom[k] = red[k][0];
}

// The sediment3D beforeEnd hooks on the DEM range [ns, ns + nd), fused per
// particle in their order (customSim.cpp:97-107): addGravity (p.f += a),
// addCustomBoundaries (the floor contact, k_dem_floor's body) and
// freezeParticles (when time <= t_freeze).
__global__ void k_hooks_dem_end(Store S, int ns, int nd, double ax, double ay, double az,
                                sphdem_params P, double tfreeze, const DevGlobals* g) {
  const int a = ns + blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns + nd || S.iam[a] != T_DEM) return;
  double4 f = S.frc[a];
  if (ax != 0.0) f.x += ax;
  if (ay != 0.0) f.y += ay;
  if (az != 0.0) f.z += az;
  const double4 p = S.posh[a];
  const double overlap = P.rmin[2] + P.dem_radius - p.z;
  if (overlap > 0.0) {
    const double od = -S.vhm[a].z;
    const double s = P.dem_k * overlap + P.dem_gamma * od;
    const double fz = s * 1.0 / S.vhm[a].w;
    S.fb[a].z += fz;
    f.z += fz;
  }
  if (g->time <= tfreeze) {
    f = make_double4(0.0, 0.0, 0.0, f.w);
    S.vel[a] = make_double4(0.0, 0.0, 0.0, S.vel[a].w);
    S.vhm[a] = make_double4(0.0, 0.0, 0.0, S.vhm[a].w);
  }
  S.frc[a] = f;
}

// ============================================================ parity helpers
// ---- calcOutputVars (sphIncompress.h:233-319): the diagnostics of an output step.
// Per-block partial sums in a fixed order, then one block adds them in block
// order.  Sums over sph (momentum, energies, counts) and over sph+sphBoundary
// (densities, maxima), as calcOutput's branches.
enum : int {
  OV_LINX = 0, OV_LINY, OV_LINZ, OV_ANG, OV_EKE, OV_EVISCF, OV_EDVISCFDT, OV_EELAST,
  OV_EDELASTDT, OV_EELASTEXACT, OV_AVEDENS, OV_DENSFROMMASS, OV_NSPH, OV_MAXV, OV_MAXF,
  OV_VARDENS, OV_COUNT
};
__global__ void __launch_bounds__(256) k_output_partials(Store S, int ns, sphdem_params P,
                                                         const double* __restrict__ densCalc,
                                                         double aveDens, int pass,
                                                         double* __restrict__ partial) {
  __shared__ double red[OV_COUNT][256];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double v[OV_COUNT];
  for (int k = 0; k < OV_COUNT; k++) v[k] = 0.0;
  if (a < ns && isSphT(S.iam[a])) {
    const double4 d = S.dpo[a];
    if (pass == 0) {
      const double4 p = S.posh[a], vh = S.vhm[a], vel = S.vel[a], f = S.frc[a];
      const double m = vh.w;
      if (S.iam[a] == T_SPH) {
        v[OV_LINX] = vel.x * m;
        v[OV_LINY] = vel.y * m;
        v[OV_LINZ] = vel.z * m;
        v[OV_ANG] = (p.x * vel.y - p.y * vel.x) * m;
        v[OV_EKE] = 0.5 * m * dot3(vh.x, vh.y, vh.z, vh.x, vh.y, vh.z);
        v[OV_EVISCF] = m * S.fvel[a].w;
        const double4 fv = S.fv[a];
        v[OV_EDVISCFDT] = m * -dot3(vh.x, vh.y, vh.z, fv.x, fv.y, fv.z);
        v[OV_NSPH] = 1.0;
      }
      v[OV_MAXV] = sqrt(dot3(vel.x, vel.y, vel.z, vel.x, vel.y, vel.z));
      v[OV_MAXF] = sqrt(dot3(f.x, f.y, f.z, f.x, f.y, f.z));
      v[OV_AVEDENS] = d.x;
      v[OV_EELAST] = m * vel.w;
      v[OV_EDELASTDT] = m * f.w;
      const double df = d.x / P.refd;
      v[OV_EELASTEXACT] = m * (P.prb * ((pow(df, 6) - 1) / 6 + 1 / df - 1)) / P.refd;
      const double dm = densCalc[a] - d.x;
      v[OV_DENSFROMMASS] = dm * dm;
    } else {
      const double dv = d.x - aveDens;
      v[OV_VARDENS] = dv * dv;
    }
  }
  for (int k = 0; k < OV_COUNT; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < OV_COUNT; k++) {
        if (k == OV_MAXV || k == OV_MAXF)
          red[k][threadIdx.x] = dmax2(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        else
          red[k][threadIdx.x] += red[k][threadIdx.x + s];
      }
    __syncthreads();
  }
  if (threadIdx.x < OV_COUNT) partial[(size_t)blockIdx.x * OV_COUNT + threadIdx.x] = red[threadIdx.x][0];
}
__global__ void k_output_final(const double* __restrict__ partial, int nblocks,
                               double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= OV_COUNT) return;
  double acc = 0.0;
  for (int b = 0; b < nblocks; b++) {
    const double x = partial[(size_t)b * OV_COUNT + k];
    acc = (k == OV_MAXV || k == OV_MAXF) ? dmax2(acc, x) : acc + x;
  }
  out[k] = acc;
}

// whole double4 streams between store order and original-index order
__global__ void k_gather_orig(const double4* __restrict__ src, const int* __restrict__ orig, int n,
                              double4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[orig[i]] = src[i];
}
__global__ void k_gather_orig_int(const int* __restrict__ src, const int* __restrict__ orig, int n,
                                  int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[orig[i]] = src[i];
}
__global__ void k_identity(int* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}
__global__ void k_get_field(const double* __restrict__ base, int off, int width,
                            const int* __restrict__ orig, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = orig[i];
  for (int c = 0; c < width; c++) out[(size_t)o * width + c] = base[(size_t)4 * i + off + c];
}
__global__ void k_set_field(double* __restrict__ base, int off, int width,
                            const int* __restrict__ orig, int n, const double* __restrict__ in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = orig[i];
  for (int c = 0; c < width; c++) base[(size_t)4 * i + off + c] = in[(size_t)o * width + c];
}
__global__ void k_get_int(const int* __restrict__ src, int add, const int* __restrict__ orig,
                          int n, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[orig[i]] = src[i] + add;
}
__global__ void k_set_int(int* __restrict__ dst, const int* __restrict__ orig, int n,
                          const int* __restrict__ in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = in[orig[i]];
}
// getCellI / dem_getCellI (dataLL.h:303-311) per original index
__global__ void k_cell_coords(Store S, int n, Grid gs, CouplingRule R, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = S.iam[i];
  double4 p = S.posh[i];
  const int o = S.orig[i];
  for (int d = 0; d < 3; d++) {
    int c;
    if (t == T_DEM)
      c = demCell(R, comp(p, d), d);
    else
      c = cellOf(comp(p, d), gs, d);
    out[3 * o + d] = c;
  }
}

// Pair enumeration of one search kind; mode 0 counts, mode 1 writes.
__global__ void k_pairs(Store S, int ns, int nd, int kind, Grid gs, Grid gc,
                        const int* __restrict__ start, Box B, CouplingRule R, int reachFull,
                        int mode, int* __restrict__ cnt, const long long* __restrict__ off,
                        int* __restrict__ oa, int* __restrict__ ob, int* __restrict__ oimg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int a;
  if (kind == SPHDEM_PAIRS_SPH || kind == SPHDEM_PAIRS_SPH2DEM) {
    if (i >= ns) return;
    a = i;
    if (!isSphT(S.iam[a])) return;
  } else {
    if (i >= nd) return;
    a = ns + i;
  }
  const double4 pa = S.posh[a];
  long long w = (mode == 1) ? off[i] : 0;
  int c = 0;
  auto emit = [&](int b, int sx, int sy, int sz) {
    if (mode == 1) {
      oa[w] = S.orig[a];
      ob[w] = S.orig[b];
      oimg[w] = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
      w++;
    }
    c++;
  };
  if (kind == SPHDEM_PAIRS_SPH || kind == SPHDEM_PAIRS_DEM2SPH) {
    forCells(gs, start, B, pa.x, pa.y, pa.z, 1, [&](int b, int sx, int sy, int sz) {
      if (b == a && (sx | sy | sz) == 0) return;
      const double4 pb = S.posh[b];
      const Img im = imagePos(pb, sx, sy, sz, B);
      const double dx = pa.x - im.x, dy = pa.y - im.y, dz = pa.z - im.z;
      const double l2 = dot3(dx, dy, dz, dx, dy, dz);
      const double lim = 2.0 * dmax2(pa.w, pb.w);
      if (l2 <= lim * lim) emit(b, sx, sy, sz);
    });
  } else if (kind == SPHDEM_PAIRS_DEM) {
    forCells(gc, start, B, pa.x, pa.y, pa.z, 1, [&](int b, int sx, int sy, int sz) {
      if (b == a && (sx | sy | sz) == 0) return;
      const double4 pb = S.posh[b];
      const Img im = imagePos(pb, sx, sy, sz, B);
      const double dx = pa.x - im.x, dy = pa.y - im.y, dz = pa.z - im.z;
      const double l2 = dot3(dx, dy, dz, dx, dy, dz);
      const double lim = 2.0 * dmax2(pa.w, pb.w);
      if (l2 <= lim * lim) emit(b, sx, sy, sz);
    });
  } else {
    const int ca[3] = {demCell(R, pa.x, 0), demCell(R, pa.y, 1), demCell(R, pa.z, 2)};
    forCells(gc, start, B, pa.x, pa.y, pa.z, reachFull, [&](int j, int sx, int sy, int sz) {
      const double4 pj = S.posh[j];
      if ((sx | sy | sz) != 0 && !demImageExists(R, B, pj, sx, sy, sz)) return;
      const Img im = imagePos(pj, sx, sy, sz, B);
      const double dx = pa.x - im.x, dy = pa.y - im.y, dz = pa.z - im.z;
      const double l2 = dot3(dx, dy, dz, dx, dy, dz);
      if (couplingAccept(R, ca, im.x, im.y, im.z, l2, true)) emit(j, sx, sy, sz);
    });
  }
  if (mode == 0) cnt[i] = c;
}

}  // namespace sphdem
