// sphdem_kernels.h -- device data layout of the gfx950 SPH-DEM loop.
//
// All particles live in one structure of arrays, kept sorted by
// (grid key, original index).  The SPH group (sph + sphBoundary) comes first,
// binned in the reference's SPH cell grid (dataLL.h:303-305); the DEM group
// follows, binned in a coarse grid whose edge covers both the DEM contact
// distance 2R and the coupling radius; deleted particles sort last.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sphdem_c.h"

namespace sphdem {

enum : int {
  T_SPH = 0,
  T_SPHB = 1,
  T_BOUNDARY = 3,
  T_DEM = 11,
  T_DELETED = 255,
};

// One particle = 12 double4 + 3 int (396 B).
//
// Periodic images are stored as ghost entries after the real particles (store
// positions [n, n + ng) after a reset, n = the uploaded count): copies of their
// origin with the position shifted by whole periods, binned in a key range of
// their own, exactly the ghosts CdataLL::updateDomain creates
// (dataLL.cpp:503-583, 631-695).
struct Store {
  double4* posh;   // r.xyz, h
  double4* vhm;    // vhat.xyz, mass
  double4* dpo;    // dens, pdr2, omega, porosity
  double4* vel;    // v.xyz, u
  double4* v0p;    // v0.xyz, press
  double4* vh0d;   // vhat0.xyz, dddt
  double4* frc;    // f.xyz, dudt
  double4* fp;     // fp.xyz, shepSum
  double4* fv;     // fv.xyz, -
  double4* fb;     // fb.xyz, -
  double4* fdrag;  // fdrag.xyz, -
  double4* fvel;   // fvel.xyz, -
  int* iam;
  int* orig;
  int* cell;       // packed cell coordinates of the particle's grid (packCell)
};
constexpr int kNumVec = 12;

// cell coordinates packed as x | y << 11 | z << 22 (x, y < 2048, z < 1024)
__host__ __device__ __forceinline__ int packCell(int x, int y, int z) {
  return x | (y << 11) | (z << 22);
}
__host__ __device__ __forceinline__ int cellX(int c) { return c & 2047; }
__host__ __device__ __forceinline__ int cellY(int c) { return (c >> 11) & 2047; }
__host__ __device__ __forceinline__ int cellZ(int c) { return (c >> 22) & 1023; }

// A run of whole cells of one (x,y) column of the SPH grid holding at most 64
// particles (a cell with more is split): the unit of work of one wavefront.
struct Segment {
  int p0, p1;  // particle range in the sorted store
  int col;     // x * dim1 + y
  int zz;      // z0 | z1 << 16
};

// Globals that the kernels read or update (the CglobalVars members of the loop).
struct DevGlobals {
  double time, dt, newDt, maxdt, oldDt;
  long long sphStep;
  double custom[9];
  unsigned long long newDt_bits;  // atomicMin target (positive doubles order as their bits)
  unsigned long long hmax_bits;   // atomicMax target
  unsigned long long ndel_sph, ndel_dem;
  unsigned int errflags;  // bit 0: a particle fell outside its cell grid
  int ghostTotal;         // ghost entries counted by the last reset's scan
};

// Cells (per axis) a ghost may occupy: one beyond the cells real particles of
// a periodic axis can hold, [cellOf(min) - 1, cellOf(max) + 1] of the domain.
struct GhostBand {
  int lo[3], hi[3];
};

// The per-particle work of the half-step just before a reset, done by the
// reset's own passes (k_wrap, k_permute<true>) instead of a launch of its
// own: mode 0 nothing, 1 start(), 2 end()'s drift of r.  Zero-initialised:
// mode 0.
struct WrapFold {
  int mode;
  int nlive;          // the half-step's range [0, nlive) of the store
  unsigned flags;     // sphdem_params::flags (CONST_H)
  double hfac;        // sphdem_params::hfac
  unsigned bodyMask;  // start(): the beforeMiddle body force's types and value
  double ax, ay, az;
};

// The per-target list search's record of an SPH-group entry or ghost: its
// position as 16-bit fixed point modulo 2^16 in the unit 1/uinv (a power of
// two) from the origin o, and the bound code B16 = ceil(((2 h / u + kQBand)^2
// + 1) / 1024) above z (the search's use of it is described with k_lists_t).
constexpr double kQBand = 1.74;
struct ListQuant {
  double o[3];  // any common origin (the grid's cmin)
  double uinv;  // 1/u, a power of two
};
__device__ __forceinline__ unsigned q16(double x, double o, double uinv) {
  return (unsigned)(long long)floor((x - o) * uinv) & 0xffffu;
}
// false (and a zero record) for a non-finite position or a non-positive h
__device__ __forceinline__ bool listRecord(const double4& p, const ListQuant& lq, uint2& r) {
  const bool ok = isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(p.w) && p.w > 0.0;
  if (!ok) {
    r = make_uint2(0u, 0u);
    return false;
  }
  const double b = 2.0 * p.w * lq.uinv + kQBand;
  const double bq = fmin(b * b + 1.0, 1.0e9);
  const unsigned b16 = (unsigned)fmin(ceil(bq * (1.0 / 1024.0)), 65535.0);
  r = make_uint2(q16(p.x, lq.o[0], lq.uinv) | (q16(p.y, lq.o[1], lq.uinv) << 16),
                 q16(p.z, lq.o[2], lq.uinv) | (b16 << 16));
  return true;
}

// Fixed-point frame of the mixed form's packed positions: grid lower corner
// and the unit.
struct Quant {
  double o[3];
  double inv;  // 1 / unit (a power of two)
  float res;   // the unit
};
__device__ __forceinline__ unsigned quant1(double x, double o, double inv) {
  const double u = rint((x - o) * inv);
  return (unsigned)fmin(fmax(u, 0.0), 4294967295.0);
}
// The mixed form's P record (fixed-point position, FP32 mass) and V record
// (FP32 vhat and h) of position-and-h p and vhat-and-mass vh.
__device__ __forceinline__ uint4 mixRecordP(const double4& p, const double4& vh, const Quant& Q) {
  return make_uint4(quant1(p.x, Q.o[0], Q.inv), quant1(p.y, Q.o[1], Q.inv),
                    quant1(p.z, Q.o[2], Q.inv), __float_as_uint((float)vh.w));
}
__device__ __forceinline__ float4 mixRecordV(const double4& p, const double4& vh) {
  return make_float4((float)vh.x, (float)vh.y, (float)vh.z, (float)p.w);
}

// Records a folded reset's permutation writes on the way (WrapFold): the
// SPH-group entries [0, ns) and the ghosts, from the values it stores.
struct PermRecords {
  int lists;  // the list search's records (rec; non-finite input to lflag[gen & 1])
  int mixed;  // the mixed form's P and V records
  int ns;
  ListQuant lq;
  uint2* rec;
  int* lflag;
  int gen;
  Quant Q;
  uint4* P;
  float4* V;
};

// Block index (xcdBlock) and global thread index (xcdThread) with the blocks
// of a 1-D grid regrouped by XCD.  The
// dispatcher deals workgroups round robin to the 8 XCDs (block b to XCD b % 8);
// here XCD x runs the x-th contiguous eighth of the blocks, in order, so the
// neighbour gathers of a cell-sorted particle range stay within that XCD's L2
// instead of every XCD streaming the whole domain.  A bijection on the grid.
__device__ __forceinline__ int xcdBlock() {
  const int nb = gridDim.x, x = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int q = nb >> 3, rem = nb & 7;
  return x * q + min(x, rem) + r;
}
__device__ __forceinline__ int xcdThread() { return xcdBlock() * blockDim.x + threadIdx.x; }

// Viscosity variant of calcViscForce (sph.h:579-631): at most one is set.
enum ViscVariant { VISC_NONE = 0, VISC_MOR = 1, VISC_CLE = 2, VISC_MON = 3 };
__host__ __device__ inline int viscVariant(uint32_t flags) {
  return (flags & SPHDEM_F_VISC_MORRIS)     ? VISC_MOR
         : (flags & SPHDEM_F_VISC_CLEARY)   ? VISC_CLE
         : (flags & SPHDEM_F_VISC_MONAGHAN) ? VISC_MON
                                            : VISC_NONE;
}
// ALPHA = viscToAlpha(VISCOSITY, H, SPSOUND) (sph.h:42, misc.h:41-43), the
// Monaghan viscosity coefficient
__host__ __device__ inline double monaghanAlpha(const sphdem_params& P) {
  return P.viscosity * (112.0 / 15.0) / (P.spsound * P.h);
}

// Uniform grid: cell_i = (int)((x_i - cmin_i)/edge + off); keys are
// base + (cx*dim1 + cy)*dim2 + cz.
struct Grid {
  double cmin[3];
  double edge;
  double off;
  int dim[3];
  int base;
  double inv;  // 1 / edge (0: every cell index takes the division)
};

struct Box {
  int periodic[3];
  double Lp[3];  // RMAX - RMIN (image shift +1)
  double Lm[3];  // RMIN - RMAX (image shift -1)
};

}  // namespace sphdem
