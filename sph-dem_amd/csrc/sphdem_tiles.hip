// sphdem_tiles.hip -- SPH neighbour search of the gfx950 loop and the pair
// passes that walk its cached lists.
//
// Search: one 64-lane wavefront per Segment (whole z-adjacent cells of one
// (x,y) column of the SPH grid, at most 64 particles; lane = target particle).
// The wavefront walks the candidate columns of the segment (the 3x3 columns
// around it, z range extended by one cell, plus periodic images), loads each
// column range 64 particles at a time into its own LDS slice (coalesced reads
// of the cell-sorted store) and every lane tests all staged candidates against
// its particle.  Per lane a candidate is admitted only if the cell of its
// (image) position lies within +-1 of the lane's cell on every axis -- the
// reference's 27-cell stencil (dataLL.cpp:1049-1071) -- and passes
// Cparticle::is_neighbr (particle.h:236-238).  Columns are walked x-outer,
// then y, then z, particles in original-index order, so each list holds the
// neighbours in the order of the reference's cached list.
//
// Passes: one thread per particle walks its list (every lane busy with an
// accepted pair, instead of the ~1/6 of lanes a divergent candidate test
// leaves active), summing in list order.
#pragma once

namespace sphdem {

constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- segments -------------------------------------------------------------------
// Greedy packing of one column's cells into segments of <= 64 particles.
template <class E>
__device__ __forceinline__ void walkColumn(const int* __restrict__ start, const Grid& g, int col,
                                           E&& emit) {
  const int base = g.base + col * g.dim[2];
  int z = 0;
  while (z < g.dim[2]) {
    const int c0 = start[base + z], c1 = start[base + z + 1];
    if (c1 == c0) {
      z++;
      continue;
    }
    if (c1 - c0 > 64) {  // an over-full cell is split into several segments
      for (int p = c0; p < c1; p += 64) emit(p, min(p + 64, c1), z, z);
      z++;
      continue;
    }
    const int z0 = z;
    int last = z, acc = c1 - c0;
    z++;
    while (z < g.dim[2]) {
      const int k = start[base + z + 1] - start[base + z];
      if (acc + k > 64) break;
      acc += k;
      if (k) last = z;
      z++;
    }
    emit(c0, start[base + last + 1], z0, last);
  }
}
__global__ void k_seg_count(const int* __restrict__ start, Grid g, int* __restrict__ cnt) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= g.dim[0] * g.dim[1]) return;
  int c = 0;
  walkColumn(start, g, col, [&](int, int, int, int) { c++; });
  cnt[col] = c;
}
__global__ void k_seg_write(const int* __restrict__ start, Grid g, const int* __restrict__ off,
                            Segment* __restrict__ segs) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= g.dim[0] * g.dim[1]) return;
  int w = off[col];
  walkColumn(start, g, col, [&](int p0, int p1, int z0, int z1) {
    segs[w++] = Segment{p0, p1, col, z0 | (z1 << 16)};
  });
}

// Real-cell range along one axis holding every particle whose image (shift s)
// can fall into the target cells [lo-1, hi+1]; a superset, clipped to the grid.
__device__ __forceinline__ void axisRange(const Grid& g, const Box& B, int ax, int s, int lo,
                                          int hi, int& r0, int& r1) {
  if (s == 0) {
    r0 = lo - 1;
    r1 = hi + 1;
  } else {
    const double L = (s == 1) ? B.Lp[ax] : B.Lm[ax];
    const double xlo = g.cmin[ax] + ((lo - 1) - g.off) * g.edge - L;
    const double xhi = g.cmin[ax] + ((hi + 2) - g.off) * g.edge - L;
    r0 = cellOf(xlo, g.cmin[ax], g.edge, g.off) - 1;
    r1 = cellOf(xhi, g.cmin[ax], g.edge, g.off) + 1;
  }
  r0 = max(r0, 0);
  r1 = min(r1, g.dim[ax] - 1);
}

// Staged candidate common part.
struct TileBase {
  double x[64], y[64], z[64];
  int cell[64];  // cell of the (image) position
  int self[64];  // index in the store, or -1 for a periodic image
};
__device__ __forceinline__ void stageBase(TileBase& T, int l, const Store& S, const Grid& g,
                                          const Box& B, int j, int sx, int sy, int sz,
                                          double4& p) {
  p = S.posh[j];
  if ((sx | sy | sz) == 0) {
    T.x[l] = p.x;
    T.y[l] = p.y;
    T.z[l] = p.z;
    T.cell[l] = S.cell[j];
    T.self[l] = j;
  } else {
    const Img im = imagePos(p, sx, sy, sz, B);
    T.x[l] = im.x;
    T.y[l] = im.y;
    T.z[l] = im.z;
    T.cell[l] = packCell(cellOf(im.x, g.cmin[0], g.edge, g.off),
                         cellOf(im.y, g.cmin[1], g.edge, g.off),
                         cellOf(im.z, g.cmin[2], g.edge, g.off));
    T.self[l] = -1;
  }
}
__device__ __forceinline__ bool inStencil(int a, int b) {
  return abs(cellX(a) - cellX(b)) <= 1 && abs(cellY(a) - cellY(b)) <= 1 &&
         abs(cellZ(a) - cellZ(b)) <= 1;
}

// The driver shared by every pass.  Consecutive workgroups land on different
// XCDs (round robin over 8), so XCD x = blockIdx % 8 gets the x-th eighth of
// the segments (a spatially coherent slab for its L2) and its wavefronts split
// that slab into contiguous runs.
template <class Op>
__global__ void __launch_bounds__(256) k_tile(Store S, Grid g, const int* __restrict__ start,
                                              Box B, const Segment* __restrict__ segs,
                                              const int* __restrict__ nsegPtr, Op op) {
  __shared__ typename Op::Tile tiles[kWavesPerBlock];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  typename Op::Tile& T = tiles[w];
  const int nseg = *nsegPtr;
  const int xcd = blockIdx.x & 7;                          // gridDim.x is a multiple of 8
  const int slab0 = (int)((long long)nseg * xcd / 8);
  const int slab1 = (int)((long long)nseg * (xcd + 1) / 8);
  const int wpx = (gridDim.x >> 3) * kWavesPerBlock;       // wavefronts per XCD
  const int lw = (blockIdx.x >> 3) * kWavesPerBlock + w;
  const int chunk = (slab1 - slab0 + wpx - 1) / wpx;
  const int s0 = slab0 + lw * chunk, s1 = min(s0 + chunk, slab1);
  typename Op::Acc acc;
  op.begin(acc);
  for (int si = s0; si < s1; si++) {
    const Segment sg = segs[si];
    const int t = sg.p0 + lane;
    const bool act = t < sg.p1;
    const int cx = sg.col / g.dim[1], cy = sg.col - cx * g.dim[1];
    const int z0 = sg.zz & 0xffff, z1 = sg.zz >> 16;
    typename Op::Lane L;
    int tcell = 0;
    if (act) {
      op.load(S, t, L);
      tcell = S.cell[t];
    }
    const int nsx = B.periodic[0] ? 3 : 1, nsy = B.periodic[1] ? 3 : 1, nsz = B.periodic[2] ? 3 : 1;
    for (int ix = 0; ix < nsx; ix++) {
      const int sx = ix == 0 ? 0 : (ix == 1 ? -1 : 1);
      int X0, X1;
      axisRange(g, B, 0, sx, cx, cx, X0, X1);
      if (X0 > X1) continue;
      for (int iy = 0; iy < nsy; iy++) {
        const int sy = iy == 0 ? 0 : (iy == 1 ? -1 : 1);
        int Y0, Y1;
        axisRange(g, B, 1, sy, cy, cy, Y0, Y1);
        if (Y0 > Y1) continue;
        for (int iz = 0; iz < nsz; iz++) {
          const int sz = iz == 0 ? 0 : (iz == 1 ? -1 : 1);
          int Z0, Z1;
          axisRange(g, B, 2, sz, z0, z1, Z0, Z1);
          if (Z0 > Z1) continue;
          for (int X = X0; X <= X1; X++)
            for (int Y = Y0; Y <= Y1; Y++) {
              const int k0 = g.base + (X * g.dim[1] + Y) * g.dim[2];
              const int j0 = start[k0 + Z0], j1 = start[k0 + Z1 + 1];
              for (int b0 = j0; b0 < j1; b0 += 64) {
                const int m = min(64, j1 - b0);
                waveSync();
                if (lane < m) op.stage(S, g, B, b0 + lane, sx, sy, sz, T, lane);
                waveSync();
                if (act) {
                  // four staged candidates per step: the tests are evaluated
                  // without branches so their LDS reads issue together, then
                  // the admitted ones are taken in order
                  for (int k = 0; k < m; k += 4) {
                    bool ok[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                      const int kk = min(k + u, 63);
                      ok[u] = (int)(k + u < m) & (int)inStencil(tcell, T.cell[kk]) &
                              (int)(T.self[kk] != t) & (int)op.test(L, T, kk);
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++)
                      if (ok[u]) op.accept(L, T, k + u);
                  }
                }
              }
            }
        }
      }
    }
    if (act) op.store(S, t, L, acc);
  }
  op.end(acc);
}

// ---- cached neighbour lists (CpInfo::neighbrs, built by calcNeighbours,
// dataLL.cpp:1026-1075, and kept until the next reset, dataLL.impl.h:186-188) --
// Entry-major layout: entry k of particle t at nbl[k * ns + t], so a wavefront
// reading entry k of 64 consecutive particles reads 256 contiguous bytes.
// An entry is the store index of the neighbour with its periodic image code in
// the top bits.
constexpr int kMaxNb = 256;
constexpr int kEntryShift = 26;  // store index < 2^26
__device__ __forceinline__ int nbIndex(int e) { return e & ((1 << kEntryShift) - 1); }
__device__ __forceinline__ int nbCode(int e) { return (int)((unsigned)e >> kEntryShift); }
__device__ __forceinline__ void codeShift(int code, int& sx, int& sy, int& sz) {
  sx = code % 3 - 1;
  sy = (code / 3) % 3 - 1;
  sz = code / 9 - 1;
}

struct ListOp {
  const double* hq;  // h to build with (null: current h)
  int ns;
  int* nbl;
  int* nbcnt;
  DevGlobals* g;
  struct Tile : TileBase {
    double hq[64];
    int ent[64];
  };
  struct Lane {
    double x, y, z, hq;
    int t, c;
  };
  struct Acc {};
  __device__ void begin(Acc&) const {}
  __device__ void end(Acc&) const {}
  __device__ void load(const Store& S, int t, Lane& L) const {
    const double4 p = S.posh[t];
    L.x = p.x;
    L.y = p.y;
    L.z = p.z;
    L.hq = hq ? hq[t] : p.w;
    L.t = t;
    L.c = 0;
  }
  __device__ void stage(const Store& S, const Grid& gg, const Box& B, int j, int sx, int sy, int sz,
                        Tile& T, int l) const {
    double4 p;
    stageBase(T, l, S, gg, B, j, sx, sy, sz, p);
    T.hq[l] = hq ? hq[j] : p.w;
    T.ent[l] = j | (((sx + 1) + 3 * (sy + 1) + 9 * (sz + 1)) << kEntryShift);
  }
  __device__ bool test(const Lane& L, const Tile& T, int k) const {
    const double dx = L.x - T.x[k], dy = L.y - T.y[k], dz = L.z - T.z[k];
    const double l2 = dot3(dx, dy, dz, dx, dy, dz);
    const double lim = 2.0 * dmax2(L.hq, T.hq[k]);
    return l2 <= lim * lim;
  }
  __device__ void accept(Lane& L, const Tile& T, int k) const {
    if (L.c < kMaxNb) nbl[(size_t)L.c * ns + L.t] = T.ent[k];
    L.c++;
  }
  __device__ void store(const Store&, int t, Lane& L, Acc&) const {
    if (L.c > kMaxNb) {
      atomicOr(&g->errflags, 2u);
      L.c = kMaxNb;
    }
    nbcnt[t] = L.c;
  }
};

// Loads neighbour entry e: store index and (image) position.
__device__ __forceinline__ void imageFix(const Box& B, int e, double4& p) {
  const int code = nbCode(e);
  if (code != 13) {
    int sx, sy, sz;
    codeShift(code, sx, sy, sz);
    const Img im = imagePos(p, sx, sy, sz, B);
    p.x = im.x;
    p.y = im.y;
    p.z = im.z;
  }
}
// Batched list walk: U entries are fetched, then their gathers are issued
// together (one memory round trip per batch instead of two per neighbour), and
// the arithmetic runs over the batch in list order.  Every neighbour in the
// SPH group is sph or sphBoundary, so no type needs to be read.
#define SPHDEM_FETCH_ENTRIES(U)                                                   \
  int e[U];                                                                       \
  _Pragma("unroll") for (int u = 0; u < U; u++) e[u] =                             \
      (k0 + u < n) ? nbl[(size_t)(k0 + u) * ns + a] : (a | (13 << kEntryShift));

// ---- calcOmega (sph.h:437-452) over the cached list -----------------------------
__global__ void __launch_bounds__(256) k_omega_l(Store S, int ns, Box B,
                                                 const int* __restrict__ nbl,
                                                 const int* __restrict__ nbcnt, double wcon) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns) return;
  const double4 pa = S.posh[a];
  const double ha = pa.w;
  double om = S.vhm[a].w * wendPdWh(wcon, 0, ha);
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    double4 p[U];
    double m[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = nbIndex(e[u]);
      p[u] = S.posh[j];
      m[u] = S.vhm[j].w;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      imageFix(B, e[u], p[u]);
      const double dx = pa.x - p[u].x, dy = pa.y - p[u].y, dz = pa.z - p[u].z;
      const double r = sqrt(dot3(dx, dy, dz, dx, dy, dz));
      om += m[u] * wendPdWh(wcon, r / ha, ha);
    }
  }
  S.dpo[a].z = 1 + ha * om / (S.dpo[a].x * 3);
}

// ---- calcDensity (sph.h:1149-1191) over the cached list -------------------------
__global__ void __launch_bounds__(256) k_density_l(Store S, int ns, Box B,
                                                   const int* __restrict__ nbl,
                                                   const int* __restrict__ nbcnt, double wcon) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns) return;
  const double4 pa = S.posh[a];
  const double ha = pa.w;
  double rho = S.vhm[a].w * wendW(wcon, 0, ha);
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    double4 p[U];
    double m[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = nbIndex(e[u]);
      p[u] = S.posh[j];
      m[u] = S.vhm[j].w;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      imageFix(B, e[u], p[u]);
      const double dx = pa.x - p[u].x, dy = pa.y - p[u].y, dz = pa.z - p[u].z;
      const double r = sqrt(dot3(dx, dy, dz, dx, dy, dz));
      rho += m[u] * wendW(wcon, r / ha, ha);
    }
  }
  S.dpo[a].x = rho;
}

// ---- calcDddtDudt (sphIncompress.h:783-851) over the cached list ----------------
__global__ void __launch_bounds__(256) k_dddt_l(Store S, int ns, Box B,
                                                const int* __restrict__ nbl,
                                                const int* __restrict__ nbcnt, double wcon,
                                                int varh) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns) return;
  const double4 pa = S.posh[a];
  const double4 va = S.vhm[a];
  const double ha = pa.w;
  const double om = S.dpo[a].z;
  double dddt = S.vh0d[a].w;
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    double4 p[U], v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = nbIndex(e[u]);
      p[u] = S.posh[j];
      v[u] = S.vhm[j];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      imageFix(B, e[u], p[u]);
      const double dx = pa.x - p[u].x, dy = pa.y - p[u].y, dz = pa.z - p[u].z;
      const double dv0 = va.x - v[u].x, dv1 = va.y - v[u].y, dv2 = va.z - v[u].z;
      const double r = sqrt(dot3(dx, dy, dz, dx, dy, dz));
      double gW[3];
      if (varh) {
        gradW(wcon, true, om, dx, dy, dz, r / ha, ha, gW);
      } else {
        double g1[3], g2[3];
        gradW(wcon, false, om, dx, dy, dz, r / ha, ha, g1);
        gradW(wcon, false, om, dx, dy, dz, r / p[u].w, p[u].w, g2);
        for (int c = 0; c < 3; c++) gW[c] = 0.5 * (g1[c] + g2[c]);
      }
      dddt += v[u].w * dot3(dv0, dv1, dv2, gW[0], gW[1], gW[2]);
    }
  }
  S.vh0d[a].w = dddt;
}

// ---- calcForce (sphIncompress.h:899-972; sph.h:579-666, 692-734) over the list --
// The per-pair dt conditions (sph.h:545-560) are min-reduced into newDt_bits.
__global__ void __launch_bounds__(256) k_force_l(Store S, int ns, Box B,
                                                 const int* __restrict__ nbl,
                                                 const int* __restrict__ nbcnt, sphdem_params P,
                                                 DevGlobals* g) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double newDt = __longlong_as_double(0x7ff0000000000000ll);
  if (a < ns) {
    const bool fluid = S.iam[a] == T_SPH;  // viscosity only acts on sph (sphIncompress.h:927)
    const double4 pa = S.posh[a];
    const double4 va = S.vhm[a];
    const double4 da = S.dpo[a];
    const double ha = pa.w;
    const bool varh = (P.flags & SPHDEM_F_VAR_H_CORRECTION) != 0;
    const bool morris = (P.flags & SPHDEM_F_VISC_MORRIS) != 0;
    const bool artif = (P.flags & SPHDEM_F_VISC_ARTIFICIAL) != 0;
    const bool liqdem = (P.flags & SPHDEM_F_LIQ_DEM) != 0;
    double4 f = S.frc[a], fpa = S.fp[a], fva = S.fv[a];
    const int n = nbcnt[a];
    constexpr int U = 4;
    for (int k0 = 0; k0 < n; k0 += U) {
      SPHDEM_FETCH_ENTRIES(U)
      double4 p[U], v[U];
      double2 d[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int j = nbIndex(e[u]);
        p[u] = S.posh[j];
        v[u] = S.vhm[j];
        d[u] = *reinterpret_cast<const double2*>(&S.dpo[j]);  // dens, pdr2
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (k0 + u >= n) break;
        imageFix(B, e[u], p[u]);
        const double hb = p[u].w, mb = v[u].w, rhob = d[u].x, pdr2b = d[u].y;
        const double dx = pa.x - p[u].x, dy = pa.y - p[u].y, dz = pa.z - p[u].z;
        const double r = sqrt(dot3(dx, dy, dz, dx, dy, dz));
        double gWa[3], gWb[3];
        if (varh) {
          gradW(P.wcon, true, da.z, dx, dy, dz, r / ha, ha, gWa);
          gradW(P.wcon, true, da.z, dx, dy, dz, r / hb, hb, gWb);
        } else {
          double g1[3], g2[3];
          gradW(P.wcon, false, da.z, dx, dy, dz, r / ha, ha, g1);
          gradW(P.wcon, false, da.z, dx, dy, dz, r / hb, hb, g2);
          for (int c = 0; c < 3; c++) gWa[c] = gWb[c] = 0.5 * (g1[c] + g2[c]);
        }
        if (fluid) {
          const double dv0 = va.x - v[u].x, dv1 = va.y - v[u].y, dv2 = va.z - v[u].z;
          const double gb0 = 0.5 * (gWa[0] + gWb[0]), gb1 = 0.5 * (gWa[1] + gWb[1]),
                       gb2 = 0.5 * (gWa[2] + gWb[2]);
          const double vdr = dot3(dx, dy, dz, dv0, dv1, dv2);
          const double hav = 0.5 * (ha + hb);
          double fv0 = 0.0, fv1 = 0.0, fv2 = 0.0;
          double vsig = 0.0;
          if (morris) {
            double visc = P.viscosity * (da.x + rhob) / (da.x * rhob);
            vsig = 2.0 * P.spsound;
            if (liqdem) visc = 2.0 * P.viscosity * P.dens / (da.x * rhob);
            const double dd = dot3(dx, dy, dz, gb0, gb1, gb2);
            const double den = r * r + 0.01 * (hav * hav);
            fv0 = dv0 * mb * visc * dd / den;
            fv1 = dv1 * mb * visc * dd / den;
            fv2 = dv2 * mb * visc * dd / den;
          }
          if (artif) {
            double avsig = 0, avisc = 0.0;
            if (vdr < 0) {
              double rr = 0.0;
              if (r != 0.0) rr = 1 / r;
              const double viss = vdr * rr;
              avsig = 2.0 * P.spsound - 4.0 * viss;
              avisc = viss * avsig * P.alpha_artificial / (da.x + rhob);
            }
            const double s = mb * avisc;
            f.x += s * gb0;
            f.y += s * gb1;
            f.z += s * gb2;
            vsig = dmax2(vsig, avsig);
          }
          newDt = dmin2(newDt, 0.6 * hav / vsig);
          newDt = dmin2(newDt, 0.125 * (hav * hav) / P.viscosity);
          fva.x += fv0;
          fva.y += fv1;
          fva.z += fv2;
          f.x += fv0;
          f.y += fv1;
          f.z += fv2;
        }
        // calcPressForce; antic = 0 under NO_ANTICLUMPING
        const double antic = 0.01 * fabs(da.y + pdr2b) * 0.0;
        double p0 = da.y * gWa[0] + pdr2b * gWb[0];
        double p1 = da.y * gWa[1] + pdr2b * gWb[1];
        double p2 = da.y * gWa[2] + pdr2b * gWb[2];
        p0 = -mb * (p0 + antic);
        p1 = -mb * (p1 + antic);
        p2 = -mb * (p2 + antic);
        fpa.x += p0;
        fpa.y += p1;
        fpa.z += p2;
        f.x += p0;
        f.y += p1;
        f.z += p2;
      }
    }
    S.frc[a] = f;
    S.fp[a] = fpa;
    S.fv[a] = fva;
  }
  atomicMinPos(&g->newDt_bits, newDt);
}

// ---- the SPH pair set (parity): count, then write (a, b, image) triples -------------
struct PairsOp {
  int mode;                // 0 count, 1 write
  int* cnt;                // per target particle
  const long long* off;    // per target particle (mode 1)
  int *oa, *ob, *oimg;
  struct Tile : TileBase {
    double h[64];
    int orig[64], code[64];
  };
  struct Lane {
    double x, y, z, h;
    int orig, c;
    long long w;
  };
  struct Acc {};
  __device__ void begin(Acc&) const {}
  __device__ void end(Acc&) const {}
  __device__ void load(const Store& S, int t, Lane& L) const {
    const double4 p = S.posh[t];
    L.x = p.x;
    L.y = p.y;
    L.z = p.z;
    L.h = p.w;
    L.orig = S.orig[t];
    L.c = 0;
    L.w = mode == 1 ? off[t] : 0;
  }
  __device__ void stage(const Store& S, const Grid& g, const Box& B, int j, int sx, int sy, int sz,
                        Tile& T, int l) const {
    double4 p;
    stageBase(T, l, S, g, B, j, sx, sy, sz, p);
    T.h[l] = p.w;
    T.orig[l] = S.orig[j];
    T.code[l] = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
  }
  __device__ bool test(const Lane& L, const Tile& T, int k) const {
    const double dx = L.x - T.x[k], dy = L.y - T.y[k], dz = L.z - T.z[k];
    const double l2 = dot3(dx, dy, dz, dx, dy, dz);
    const double lim = 2.0 * dmax2(L.h, T.h[k]);
    return l2 <= lim * lim;
  }
  __device__ void accept(Lane& L, const Tile& T, int k) const {
    if (mode == 1) {
      oa[L.w] = L.orig;
      ob[L.w] = T.orig[k];
      oimg[L.w] = T.code[k];
      L.w++;
    }
    L.c++;
  }
  __device__ void store(const Store&, int t, Lane& L, Acc&) const {
    if (mode == 0) cnt[t] = L.c;
  }
};

}  // namespace sphdem
