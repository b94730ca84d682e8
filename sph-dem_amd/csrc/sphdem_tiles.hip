// sphdem_tiles.hip -- SPH neighbour search of the gfx950 loop and the pair
// passes that walk its cached lists.
//
// Search: one 64-lane wavefront per Segment (whole z-adjacent cells of one
// (x,y) column of the SPH grid, at most 64 particles; lane = target particle).
// The wavefront streams the candidates of the 3x3 columns around its segment
// (z range extended by one cell) from the cell-sorted store -- real particles
// and the periodic ghosts, which live in cells of their own key range -- with
// wave-uniform scalar loads; every lane tests every candidate against its
// particle.  Columns are walked x-outer, then y, then z, and within a cell
// the real particles (original-index order) precede the ghosts (receive
// order), so each list holds the neighbours in the order of the reference's
// cached list (dataLL.cpp:1049-1071).
//
// Passes: one thread per particle walks its list (every lane busy with an
// accepted pair), summing in list order.
#pragma once

namespace sphdem {

constexpr int kWavesPerBlock = 4;

// ---- segments -------------------------------------------------------------------
// Greedy packing of one column's cells into segments of <= cap particles
// (cap <= 64, the tile shape's segment size); s[z] = store start of cell z of
// the column, z = 0..Z (s[Z] the column end).
template <class E>
__device__ __forceinline__ void walkColumn(const int* s, int Z, int cap, E&& emit) {
  int z = 0;
  while (z < Z) {
    const int c0 = s[z], c1 = s[z + 1];
    if (c1 == c0) {
      z++;
      continue;
    }
    if (c1 - c0 > cap) {  // an over-full cell is split into several segments
      for (int p = c0; p < c1; p += cap) emit(p, min(p + cap, c1), z, z);
      z++;
      continue;
    }
    const int z0 = z;
    int last = z, acc = c1 - c0;
    z++;
    while (z < Z) {
      const int k = s[z + 1] - s[z];
      if (acc + k > cap) break;
      acc += k;
      if (k) last = z;
      z++;
    }
    emit(c0, s[last + 1], z0, last);
  }
}
// One wavefront per column: the column's Z + 1 cell starts are staged in LDS,
// then every lane runs the (wave-uniform) packing and lane 0 writes.
constexpr int kSegWaves = 4;
constexpr int kMaxZ = 1024;  // the grid's z limit
template <class E>
__device__ __forceinline__ bool stageColumn(const int* __restrict__ start, const Grid& g,
                                            int (*lds)[kMaxZ + 1], int& col, E&& body) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  col = blockIdx.x * kSegWaves + w;
  if (col >= g.dim[0] * g.dim[1]) return false;
  const int Z = g.dim[2];
  const int* s = start + g.base + col * Z;
  for (int z = lane; z <= Z; z += 64) lds[w][z] = s[z];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  body(lds[w], Z, lane);
  return true;
}
__global__ void __launch_bounds__(64 * kSegWaves) k_seg_count(const int* __restrict__ start,
                                                               Grid g, int cap,
                                                               int* __restrict__ cnt) {
  __shared__ int lds[kSegWaves][kMaxZ + 1];
  int col;
  const int ncol = g.dim[0] * g.dim[1];
  if (!stageColumn(start, g, lds, col, [&](const int* s, int Z, int lane) {
        int c = 0;
        walkColumn(s, Z, cap, [&](int, int, int, int) { c++; });
        if (lane == 0) cnt[col] = c;
      }))
    if (col == ncol && (threadIdx.x & 63) == 0) cnt[ncol] = 0;  // the scan's closing zero
}
__global__ void __launch_bounds__(64 * kSegWaves) k_seg_write(const int* __restrict__ start,
                                                               Grid g, int cap,
                                                               const int* __restrict__ off,
                                                               Segment* __restrict__ segs) {
  __shared__ int lds[kSegWaves][kMaxZ + 1];
  int col;
  stageColumn(start, g, lds, col, [&](const int* s, int Z, int lane) {
    int w = off[col];
    walkColumn(s, Z, cap, [&](int p0, int p1, int z0, int z1) {
      if (lane == 0) segs[w] = Segment{p0, p1, col, z0 | (z1 << 16)};
      w++;
    });
  });
}

// Segments [s0, s1) of wavefront w of this workgroup.  Consecutive workgroups
// land on different XCDs (round robin over 8), so XCD x = blockIdx % 8 gets the
// x-th eighth of the segments (a spatially coherent slab for its L2) and its
// wavefronts split that slab into contiguous runs.  gridDim.x is a multiple of 8.
__device__ __forceinline__ void segRange(int nseg, int w, int& s0, int& s1) {
  const int xcd = blockIdx.x & 7;
  const int slab0 = (int)((long long)nseg * xcd / 8);
  const int slab1 = (int)((long long)nseg * (xcd + 1) / 8);
  const int wpx = (gridDim.x >> 3) * kWavesPerBlock;  // wavefronts per XCD
  const int lw = (blockIdx.x >> 3) * kWavesPerBlock + w;
  const int chunk = (slab1 - slab0 + wpx - 1) / wpx;
  s0 = slab0 + lw * chunk;
  s1 = min(s0 + chunk, slab1);
}

// ---- cached neighbour lists (CpInfo::neighbrs, built by calcNeighbours,
// dataLL.cpp:1026-1075, and kept until the next reset, dataLL.impl.h:186-188) --
// Entries are grouped in eights per particle, group-major: entry k of particle
// t is nbl[((k / 8) * ns + t) * 8 + k % 8], so a wavefront reading the next
// four entries of 64 consecutive particles reads 64 16-byte pieces of one
// 2 KiB block, and the search writes each 32-byte group with two 16-byte
// stores of one lane, close in time (whole 32-byte sectors reach memory).  An
// entry is the store index of the neighbour (a ghost's index for a periodic
// image).
constexpr int kMaxNb = 256;  // a multiple of the group size
constexpr int kChunk = 4;    // entries per store
constexpr int kGroup = 8;    // entries per 32-byte group
__device__ __forceinline__ size_t nbEntry(int k, int t, int ns) {
  return ((size_t)(k >> 3) * ns + t) * kGroup + (k & 7);
}
// Pointer into the constant address space (scalar loads for uniform indices).
template <class T>
using CPtr = const __attribute__((address_space(4))) T*;

// bits [0, k) set, k in [0, 32]
__device__ __forceinline__ unsigned lowBits(int k) { return k >= 32 ? ~0u : (1u << k) - 1u; }

// calcNeighbours for every real SPH-group particle, one wavefront per Segment.
//
// Per candidate column (X, Y): the real cells Z0..Z1 hold store range
// [start[k0+Z0], start[k0+Z1+1]), the ghost cells the same range of the ghost
// keys.  Without ghosts the whole range is one stream and the 27-cell stencil
// reduces to "j lies in the store range of cells cz-1..cz+1", an integer range
// per lane; with ghosts the column is streamed cell by cell (the real part of
// a cell, then its ghosts) with the lane's |z - cz| <= 1 test.  is_neighbr
// (particle.h:236-238) is len2(ra - rb) <= (2 max(ha, hb))^2 with blitz's dot
// association.  Candidates arrive by scalar loads, two per load, three loads
// ahead of the arithmetic.
__global__ void __launch_bounds__(256) k_lists(const double4* __restrict__ pos,
                                               const int* __restrict__ cellc, Grid g, int gbase,
                                               const int* __restrict__ start,
                                               const Segment* __restrict__ segs,
                                               const int* __restrict__ nsegPtr, int ns,
                                               int* __restrict__ nbl, int* __restrict__ nbcnt,
                                               DevGlobals* dg) {
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // Every address of the candidate stream is wave-uniform and nothing here
  // writes the store, the cell starts or the segments, so they are read through
  // the constant address space: scalar loads, one per candidate per wavefront.
  const CPtr<double> cpos = (CPtr<double>)pos;
  const CPtr<int> cstart = (CPtr<int>)start;
  const CPtr<int> csegs = (CPtr<int>)segs;
  int s0, s1;
  segRange(*(CPtr<int>)nsegPtr, w, s0, s1);
  for (int si = s0; si < s1; si++) {
    const Segment sg{csegs[4 * si], csegs[4 * si + 1], csegs[4 * si + 2], csegs[4 * si + 3]};
    const int t = sg.p0 + lane;
    const bool act = t < sg.p1;
    const int cx = sg.col / g.dim[1], cy = sg.col - cx * g.dim[1];
    const int z0 = sg.zz & 0xffff, z1 = sg.zz >> 16;
    const int Z0 = max(z0 - 1, 0), Z1 = min(z1 + 1, g.dim[2] - 1);
    double px = 0.0, py = 0.0, pz = 0.0, lima2 = 0.0;
    int cz = -4;  // idle lanes: no cell is within one of it
    if (act) {
      const double4 p = pos[t];
      px = p.x;
      py = p.y;
      pz = p.z;
      lima2 = (2.0 * p.w) * (2.0 * p.w);
      cz = cellZ(cellc[t]);
    }
    int c = 0;  // entries found
    // the chunk being filled, as a shift register: b0 the newest entry
    int b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    auto push = [&](int e) {
      b3 = b2;
      b2 = b1;
      b1 = b0;
      b0 = e;
      if ((c & (kChunk - 1)) == kChunk - 1 && c < kMaxNb)
        *reinterpret_cast<int4*>(nbl + nbEntry(c & ~3, t, ns)) = make_int4(b3, b2, b1, b0);
      c++;
    };
    // candidates [J0, J1), admitted for a lane when j lies in [lo, lo + wd) and
    // is not the lane's own particle.  Batches of 32: the distance tests fill a
    // per-lane bit mask (candidate jb + k in bit 31 - k), the lane's index
    // window is a second mask, and then each lane appends its accepted
    // candidates in order.  Candidates come two at a time (one 64-byte scalar
    // read; loads past J1 are clamped to J1 - 1, whose two entries the store's
    // slack covers) into a ring of four groups, three groups ahead of the test.
    auto stream = [&](int J0, int J1, int lo, int wd) {
      for (int jb = J0; jb < J1; jb += 32) {
        const int klo = min(max(lo - jb, 0), 32), khi = min(max(lo + wd - jb, 0), 32);
        unsigned win = lowBits(khi) & ~lowBits(klo);
        const int self = t - jb;
        if ((unsigned)self < 32u) win &= ~(1u << self);
        win = __builtin_bitreverse32(win);
        unsigned acc = 0;
        double4 q[4][2];  // ring of four groups of two candidates
        auto fetch = [&](double4* d, int j) {
          const CPtr<double> b = cpos + 4 * min(j, J1 - 1);
#pragma unroll
          for (int u = 0; u < 2; u++)
            d[u] = make_double4(b[4 * u], b[4 * u + 1], b[4 * u + 2], b[4 * u + 3]);
        };
        auto test = [&](const double4* d) {
#pragma unroll
          for (int u = 0; u < 2; u++) {
            const double dx = px - d[u].x, dy = py - d[u].y, dz = pz - d[u].z;
            const double l2 = dot3(dx, dy, dz, dx, dy, dz);
            const double hb2 =
                __longlong_as_double(__double_as_longlong(d[u].w) + (1ll << 52));
            acc = (acc << 1) | (l2 <= fmax(lima2, hb2 * hb2) ? 1u : 0u);
          }
        };
        fetch(q[0], jb);
        fetch(q[1], jb + 2);
        fetch(q[2], jb + 4);
#pragma unroll
        for (int gq = 0; gq < 16; gq++) {
          if (gq + 3 < 16) fetch(q[(gq + 3) & 3], jb + 2 * (gq + 3));
          test(q[gq & 3]);
        }
        unsigned m = acc & win;
        int j = jb;
        while (m) {
          const int z = __builtin_clz(m);
          j += z;
          push(j);
          j++;
          m = (m << z) << 1;
        }
      }
    };
    for (int X = max(cx - 1, 0); X <= min(cx + 1, g.dim[0] - 1); X++)
      for (int Y = max(cy - 1, 0); Y <= min(cy + 1, g.dim[1] - 1); Y++) {
        const int col = (X * g.dim[1] + Y) * g.dim[2];
        const int k0 = g.base + col, q0 = gbase + col;
        const int J0 = cstart[k0 + Z0], J1 = cstart[k0 + Z1 + 1];
        const bool ghosts = gbase > 0 && cstart[q0 + Z0] != cstart[q0 + Z1 + 1];
        if (!ghosts) {
          if (J0 >= J1) continue;
          // store range of this column's cells cz-1 .. cz+1 (empty for idle lanes)
          int lo = 0, wd = 0;
          if (act) {
            lo = start[k0 + max(cz - 1, 0)];
            wd = start[k0 + min(cz + 2, g.dim[2])] - lo;
          }
          stream(J0, J1, lo, wd);
        } else {
          for (int z = Z0; z <= Z1; z++) {
            const bool near = abs(z - cz) <= 1;
            const int r0 = cstart[k0 + z], r1 = cstart[k0 + z + 1];
            if (r0 < r1) stream(r0, r1, r0, near ? r1 - r0 : 0);
            const int h0 = cstart[q0 + z], h1 = cstart[q0 + z + 1];
            if (h0 < h1) stream(h0, h1, h0, near ? h1 - h0 : 0);
          }
        }
      }
    if (act) {
      if (c > kMaxNb) {
        atomicOr(&dg->errflags, 2u);
        c = kMaxNb;
      }
      const int rem = c & (kChunk - 1);  // entries of the last, partial chunk
      if (rem)
        *reinterpret_cast<int4*>(nbl + nbEntry(c & ~3, t, ns)) =
            rem == 1 ? make_int4(b0, 0, 0, 0)
                     : (rem == 2 ? make_int4(b1, b0, 0, 0) : make_int4(b2, b1, b0, 0));
      nbcnt[t] = c;
    }
  }
}

// ---- per-target search -------------------------------------------------------------
// The same lists from one lane per target particle.  A lane walks its own
// 27-cell stencil (x-outer, then y, then z; per column the cells cz-1..cz+1 as
// one store range, or, in a column holding ghosts, cell by cell: the real part,
// then the ghosts), so each lane tests only its own ~27 cells of candidates
// instead of the union a segment's wavefront streams.  Lanes of one wavefront
// are 64 consecutive targets of the cell-sorted store: they share most of their
// stencil and read the same candidates at the same time.
//
// The test runs on quantised records first: a candidate's position as 16-bit
// fixed point modulo 2^16 in a unit u (a power of two with edge/u in
// [1024, 2048)), so a stencil separation (< 2 cell edges per axis) is an exact
// int16 difference of at most 4097 units, and its squared length D an exact
// int32 (one packed 16-bit subtraction, one 16-bit dot product and one 16-bit
// multiply-add).  Quantisation moves |dr| by less than sqrt(3) u, so with
// band = 1.74 > sqrt(3):
//   D <= (2 h_a / u - band)^2            is_neighbr holds (2 max(h) >= 2 h_a);
//   D >  (2 h / u + band)^2 for h_a, h_b is_neighbr fails;
// anything between is decided by the FP64 test on the store positions, in the
// reference's arithmetic.  Each bound keeps a margin of more than 0.007 u,
// far above the FP64 rounding of |dr|^2 (about 1e-15 relative), so the list is
// the FP64 test's exactly.  The record's upper half carries the candidate's
// bound as B16 = ceil(((2 h_b / u + band)^2 + 1) / 1024) in bits 16..31 above
// its z: the word shifted right by 6 is at least 1024 * B16, so D at most that
// is "maybe" whenever D <= 1024 * B16.  A non-finite position
// or a non-positive h anywhere sets a flag that sends every candidate to the
// FP64 test.
// (kQBand, ListQuant, q16 and listRecord: sphdem_kernels.h)
// Records of the SPH-group entries [0, ns) and the ghosts [n, n + ng).  The
// flag slot of generation gen collects non-finite input; the next
// generation's slot is cleared for the next build.
__global__ void k_lrec(const double4* __restrict__ posh, int ns, int n, int ng, ListQuant lq,
                       uint2* __restrict__ rec, int* __restrict__ lflag, int gen,
                       const int* __restrict__ ngp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) lflag[(gen + 1) & 1] = 0;
  if (i >= ns + (ngp ? min(*ngp, ng) : ng)) return;
  const int s = i < ns ? i : n + (i - ns);
  uint2 r;
  if (!listRecord(posh[s], lq, r)) atomicOr(&lflag[gen & 1], 1);
  rec[s] = r;
}

// acc = 2 acc + (x >= D ? 1 : 0): a compare into VCC and an add with carry,
// where (acc << 1) | (x >= D) compiles to a compare, a select and a merge.
// After NB steps candidate k of a batch is bit NB - 1 - k.
__device__ __forceinline__ void shiftInGe(unsigned& acc, int x, int D) {
  asm("v_cmp_ge_i32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
      : "+v"(acc)
      : "v"(x), "v"(D)
      : "vcc");
}

typedef short short2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ short2v asShort2(unsigned v) { return __builtin_bit_cast(short2v, v); }

// Grid: 8 * ceil(ceil(ns / 8) / blockDim) workgroups; XCD blockIdx % 8 takes
// the corresponding eighth of the targets (a spatially coherent slab for its
// L2).  NB = 32 candidates per batch (the masks' width).
template <int NB>
__global__ void __launch_bounds__(256) k_lists_t(const double4* __restrict__ pos,
                                                 const uint2* __restrict__ rec,
                                                 const int* __restrict__ cellc, Grid g, int gbase,
                                                 const int* __restrict__ start, int ns,
                                                 double uinv, const int* __restrict__ lflag,
                                                 int gen, int* __restrict__ nbl,
                                                 int* __restrict__ nbcnt, DevGlobals* dg) {
  static_assert(NB == 32, "the masks are 32 bits wide");
  const int per = (ns + 7) >> 3;
  const int local = (int)(blockIdx.x >> 3) * (int)blockDim.x + (int)threadIdx.x;
  const int t = (int)(blockIdx.x & 7) * per + local;
  if (local >= per || t >= ns) return;
  const double4 p = pos[t];
  const uint2 own = rec[t];
  const short2v ownxy = asShort2(own.x);
  int aq = -1, bq = 0x7fffffff;
  if (lflag[gen & 1] == 0) {
    const double a = 2.0 * p.w * uinv - kQBand;
    if (a > 0.0) aq = (int)floor(a * a) - 1;
    const double b = 2.0 * p.w * uinv + kQBand;
    bq = (int)fmin(ceil(b * b) + 1.0, 2.0e9);
  }
  const double lima2 = (2.0 * p.w) * (2.0 * p.w);
  const int cell = cellc[t];
  const int cx = cell & 2047, cy = (cell >> 11) & 2047, cz = cellZ(cell);
  int c = 0;  // entries found
  int b0 = 0, b1 = 0, b2 = 0, b3 = 0;  // the chunk being filled, b0 the newest
  auto push = [&](int e) {
    b3 = b2;
    b2 = b1;
    b1 = b0;
    b0 = e;
    if ((c & (kChunk - 1)) == kChunk - 1 && c < kMaxNb)
      *reinterpret_cast<int4*>(nbl + nbEntry(c & ~3, t, ns)) = make_int4(b3, b2, b1, b0);
    c++;
  };
  // candidates [J0, J1) in batches of NB: bit k of acc / may for candidate
  // jb + k (reads past J1 land in the records' slack and are masked off)
  auto scan = [&](int J0, int J1) {
    for (int jb = J0; jb < J1; jb += NB) {
      const uint2* rp = rec + jb;
      uint2 r[NB];
#pragma unroll
      for (int k = 0; k < NB; k++) r[k] = rp[k];
      unsigned acc = 0, may = 0;
#pragma unroll
      for (int k = 0; k < NB; k++) {
        const short2v dxy = asShort2(r[k].x) - ownxy;
        const int dz = (int)(short)(r[k].y - own.y);
        const int D = __builtin_amdgcn_sdot2(dxy, dxy, dz * dz, false);
        shiftInGe(acc, aq, D);
        shiftInGe(may, max(bq, (int)(r[k].y >> 6)), D);
      }
      // candidate jb + k is bit 31 - k
      unsigned win = __builtin_bitreverse32(lowBits(min(J1 - jb, NB)));
      const int self = t - jb;
      if ((unsigned)self < (unsigned)NB) win &= ~(0x80000000u >> self);
      unsigned amb = may & ~acc & win;
      while (amb) {  // the FP64 test of is_neighbr (particle.h:236-238)
        const int k = __builtin_clz(amb);
        amb &= ~(0x80000000u >> k);
        const double4 q = pos[jb + k];
        const double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const double l2 = dot3(dx, dy, dz, dx, dy, dz);
        const double hb2 = __longlong_as_double(__double_as_longlong(q.w) + (1ll << 52));
        if (l2 <= fmax(lima2, hb2 * hb2)) acc |= 0x80000000u >> k;
      }
      unsigned m = acc & win;
      while (m) {
        const int k = __builtin_clz(m);
        m &= ~(0x80000000u >> k);
        push(jb + k);
      }
    }
  };
  const int z0 = max(cz - 1, 0), z1 = min(cz + 1, g.dim[2] - 1);
  for (int X = max(cx - 1, 0); X <= min(cx + 1, g.dim[0] - 1); X++)
    for (int Y = max(cy - 1, 0); Y <= min(cy + 1, g.dim[1] - 1); Y++) {
      const int col = (X * g.dim[1] + Y) * g.dim[2];
      const int k0 = g.base + col, q0 = gbase + col;
      if (gbase > 0 && start[q0 + z0] != start[q0 + z1 + 1]) {
        for (int z = z0; z <= z1; z++) {
          scan(start[k0 + z], start[k0 + z + 1]);
          scan(start[q0 + z], start[q0 + z + 1]);
        }
      } else {
        scan(start[k0 + z0], start[k0 + z1 + 1]);
      }
    }
  if (c > kMaxNb) {
    atomicOr(&dg->errflags, 2u);
    c = kMaxNb;
  }
  const int rem = c & (kChunk - 1);  // entries of the last, partial chunk
  if (rem)
    *reinterpret_cast<int4*>(nbl + nbEntry(c & ~3, t, ns)) =
        rem == 1 ? make_int4(b0, 0, 0, 0)
                 : (rem == 2 ? make_int4(b1, b0, 0, 0) : make_int4(b2, b1, b0, 0));
  nbcnt[t] = c;
}

// The cached lists as (a, b, image) triples of original indices, for the
// parity enumeration of the SPH pair set: a ghost neighbour reports its
// origin and its image code.
__global__ void k_list_pairs(Store S, int ns, int n, const int* __restrict__ nbl,
                             const int* __restrict__ nbcnt, const int* __restrict__ gcode,
                             const long long* __restrict__ off, int* __restrict__ oa,
                             int* __restrict__ ob, int* __restrict__ oimg) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns) return;
  const int cnt = nbcnt[a];
  const int ia = S.orig[a];
  long long w = off[a];
  for (int k = 0; k < cnt; k++, w++) {
    const int e = nbl[nbEntry(k, a, ns)];
    oa[w] = ia;
    ob[w] = S.orig[e];
    oimg[w] = e >= n ? gcode[e - n] : 13;
  }
}

// Batched list walk: U entries are fetched, then their gathers are issued
// together (one memory round trip per batch instead of two per neighbour), and
// the arithmetic runs over the batch in list order.  Every neighbour in the
// SPH group is sph or sphBoundary (or a ghost of one, already at its image
// position), so no type needs to be read.
#define SPHDEM_FETCH_ENTRIES(U)                                                    \
  int e[U];                                                                        \
  _Pragma("unroll") for (int v = 0; v < U / kChunk; v++) {                         \
    const int4 ch = *reinterpret_cast<const int4*>(nbl + nbEntry(k0 + 4 * v, a, ns)); \
    const int ce[4] = {ch.x, ch.y, ch.z, ch.w};                                      \
    _Pragma("unroll") for (int i = 0; i < kChunk; i++) {                           \
      const int u = v * kChunk + i;                                                \
      e[u] = (k0 + u < n) ? ce[i] : a;                                             \
    }                                                                              \
  }

// ---- per-pass reciprocals --------------------------------------------------------
// Quotients by a value that serves many pairs (the particle's own h and omega,
// a pair's viscous denominator for its three components) are formed with divr
// from one correctly rounded reciprocal: bit for bit the IEEE quotient the
// reference computes, at a fraction of a division's cost.  Per neighbour, the
// pair passes gather hinv = {1/h, 1/h^5} of the neighbour (computed once per
// particle by k_hinv whenever h may have changed).
__global__ void k_hinv(const double4* __restrict__ posh, int ns, double2* __restrict__ hinv) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ns) return;
  const double h = posh[a].w;
  hinv[a] = make_double2(1.0 / h, 1.0 / pow5(h));
}


// {x, y, z, mass} of every SPH-group entry [0, ns) and ghost [n, n + ng):
// what the omega and density sums read of a neighbour, in one 32-byte record
// (two loads instead of three).
__global__ void k_pack_posm(Store S, int ns, int n, int ng, double4* __restrict__ posm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns + ng) return;
  const int e = i < ns ? i : n + (i - ns);
  const double4 p = S.posh[e];
  posm[e] = make_double4(p.x, p.y, p.z, S.vhm[e].w);
}

// ---- calcOmega (sph.h:437-452) over the cached list -----------------------------
__global__ void __launch_bounds__(256) k_omega_l(Store S, int ns, Box B,
                                                 const int* __restrict__ nbl,
                                                 const int* __restrict__ nbcnt,
                                                 const double4* __restrict__ posm, double wcon) {
  const int a = xcdThread();
  if (a >= ns) return;
  const double4 pa = S.posh[a];
  const double ha = pa.w;
  const double rha = 1.0 / ha, i3 = 1.0 / cube(ha), c3 = 3.0 / pow4(ha);
  double om = S.vhm[a].w * wendPdWhi(wcon, 0, ha, rha, i3, c3);
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    double4 p[U];  // x, y, z, mass
#pragma unroll
    for (int u = 0; u < U; u++) p[u] = posm[e[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      const double dx = pa.x - p[u].x, dy = pa.y - p[u].y, dz = pa.z - p[u].z;
      const double r = sqrt(dot3(dx, dy, dz, dx, dy, dz));
      om += p[u].w * wendPdWhi(wcon, divr(r, ha, rha), ha, rha, i3, c3);
    }
  }
  S.dpo[a].z = 1 + ha * om / (S.dpo[a].x * 3);
}

// ---- calcDensity (sph.h:1149-1191) over the cached list -------------------------
// out: where the sums go (null: the particles' dens)
__global__ void __launch_bounds__(256) k_density_l(Store S, int ns, Box B,
                                                   const int* __restrict__ nbl,
                                                   const int* __restrict__ nbcnt,
                                                   const double4* __restrict__ posm, double wcon,
                                                   double* __restrict__ out = nullptr) {
  const int a = xcdThread();
  if (a >= ns) return;
  const double4 pa = S.posh[a];
  const double ha = pa.w;
  const double rha = 1.0 / ha, i3 = 1.0 / cube(ha);
  double rho = S.vhm[a].w * wendWi(wcon, 0, i3);
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    double4 p[U];  // x, y, z, mass
#pragma unroll
    for (int u = 0; u < U; u++) p[u] = posm[e[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      const double dx = pa.x - p[u].x, dy = pa.y - p[u].y, dz = pa.z - p[u].z;
      const double r = sqrt(dot3(dx, dy, dz, dx, dy, dz));
      rho += p[u].w * wendWi(wcon, divr(r, ha, rha), i3);
    }
  }
  if (out)
    out[a] = rho;
  else
    S.dpo[a].x = rho;
}

// ---- calcDddtDudt (sphIncompress.h:783-851) over the cached list ----------------
// OMG (VAR_H only): calcOmega fused into the walk, as in k_force_l -- every
// dddt term of a is linear in 1/omega_a, so the terms are summed without it
// beside the omega sum and scaled once.  Three wavefronts per SIMD: 168 VGPRs
// without spills (170 would allow two), 2 % faster.
template <bool VARH, bool OMG = false>
__global__ void __launch_bounds__(256, 3) k_dddt_l(Store S, int ns, Box B,
                                                const int* __restrict__ nbl,
                                                const int* __restrict__ nbcnt,
                                                const double2* __restrict__ hinv, double wcon) {
  static_assert(VARH || !OMG, "omega is only fused under VAR_H");
  const int a = xcdThread();
  if (a >= ns) return;
  const double4 pa = S.posh[a];
  const double4 va = S.vhm[a];
  const double ha = pa.w;
  const double rha = 1.0 / ha;
  const double rom = (VARH && !OMG) ? 1.0 / S.dpo[a].z : 1.0;
  const double i3 = OMG ? 1.0 / cube(ha) : 0.0, c3 = OMG ? 3.0 / pow4(ha) : 0.0;
  double om = OMG ? va.w * wendPdWhi(wcon, 0, ha, rha, i3, c3) : 0.0;
  // gradW = dr * G with G = poly(q) K(h) / r, K(h) = WCON h^-4 (/ omega_a
  // under VAR_H, sph.h:331-332): F(q, h) = WCON poly(q) / (q h^5) and 1/q = h/r
  const double Ka = wcon * (rha * rha) * (rha * rha) * rom;
  double dddt = OMG ? 0.0 : S.vh0d[a].w;
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    double4 p[U], v[U];
    double2 hi[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = e[u];
      p[u] = S.posh[j];
      v[u] = S.vhm[j];
      if (!VARH) hi[u] = hinv[j];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      const double dx = pa.x - p[u].x, dy = pa.y - p[u].y, dz = pa.z - p[u].z;
      const double dv0 = va.x - v[u].x, dv1 = va.y - v[u].y, dv2 = va.z - v[u].z;
      const double r = sqrt(dot3(dx, dy, dz, dx, dy, dz));
      const double rinv = r > 0.0 ? 1.0 / r : 0.0;  // r = 0: dr = 0, gradW = 0
      const double qa = divr(r, ha, rha);
      if (OMG) om += v[u].w * wendPdWhi(wcon, qa, ha, rha, i3, c3);
      double G = qa <= 2.0 ? wendPoly(qa) * Ka * rinv : 0.0;
      if (!VARH) {
        const double qb = divr(r, p[u].w, hi[u].x);
        const double Kb = wcon * (hi[u].y * p[u].w);  // WCON hb^-4
        const double Gb = qb <= 2.0 ? wendPoly(qb) * Kb * rinv : 0.0;
        G = 0.5 * (G + Gb);  // the average gradient
      }
      dddt += v[u].w * (dot3(dv0, dv1, dv2, dx, dy, dz) * G);
    }
  }
  if (OMG) {
    const double omega = 1 + ha * om / (S.dpo[a].x * 3);
    S.dpo[a].z = omega;
    dddt = S.vh0d[a].w + (1.0 / omega) * dddt;
  }
  S.vh0d[a].w = dddt;
}

// ---- calcForce (sphIncompress.h:899-972; sph.h:579-666, 692-734) over the list --
// One instantiation per combination of the flags that change the pair
// arithmetic (VISC: the viscosity variant, VISC_NONE..VISC_MON).  The per-pair
// dt conditions (sph.h:545-560) are min-reduced into newDt_bits.  When the
// signal speed is the constant 2*SPSOUND (or 0) -- Morris or Cleary without
// artificial viscosity -- x -> RN(x/c) is monotone, so the minimum over pairs
// of RN(0.6*hav / vsig) is RN(min(0.6*hav) / vsig): the numerators are
// min-reduced and divided once (likewise c_diff*hav^2 / VISCOSITY).
//
// OMG (VAR_H only): calcOmega fused into the same walk.  Under VAR_H every
// force term of a is linear in 1/omega_a and in nothing else of omega
// (sph.h:331-332), and omega_a is a sum over a's own list (sph.h:437-452),
// so the walk sums the omega terms (in list order, as k_omega_l) beside the
// force terms taken without 1/omega_a, stores omega_a, and adds the force
// sums times 1/omega_a to the incoming f, fp, fv.
template <bool VARH, int VISC, bool ARTIF, bool LIQDEM, bool OMG = false>
__global__ void __launch_bounds__(256) k_force_l(Store S, int ns, Box B,
                                                 const int* __restrict__ nbl,
                                                 const int* __restrict__ nbcnt,
                                                 const double2* __restrict__ hinv,
                                                 sphdem_params P, DevGlobals* g) {
  static_assert(VARH || !OMG, "omega is only fused under VAR_H");
  const int a = xcdThread();
  const double inf = __longlong_as_double(0x7ff0000000000000ll);
  double newDt = inf;
  if (a < ns) {
    const bool fluid = S.iam[a] == T_SPH;  // viscosity only acts on sph (sphIncompress.h:927)
    const double4 pa = S.posh[a];
    const double4 va = S.vhm[a];
    const double4 da = S.dpo[a];
    const double ha = pa.w;
    const double rha = 1.0 / ha;
    const double rom = (VARH && !OMG) ? 1.0 / da.z : 1.0;
    const double i3 = OMG ? 1.0 / cube(ha) : 0.0, c3 = OMG ? 3.0 / pow4(ha) : 0.0;
    double om = OMG ? va.w * wendPdWhi(P.wcon, 0, ha, rha, i3, c3) : 0.0;
    // gradW = dr * F(q, h) (/ omega_a under VAR_H), F = WCON poly(q) / (q h^5)
    // = dr * poly(q) * K(h) / r with K(h) = WCON h^-4 (/ omega_a)
    const double Ka = P.wcon * (rha * rha) * (rha * rha) * rom;
    const double wrom = P.wcon * rom;
    const double viscLiq = 2.0 * P.viscosity * P.dens;  // numerator under LIQ_DEM
    const double vsigC = (VISC == VISC_MOR || VISC == VISC_CLE) ? 2.0 * P.spsound : 0.0;
    const double cdiff = VISC == VISC_MON ? 0.06 : 0.125;  // viscDiffusionCondition
    const double alphaM = monaghanAlpha(P);
    double minA = inf, minB = inf;  // min over pairs of 0.6*hav and cdiff*hav^2
    const double4 zero4 = make_double4(0.0, 0.0, 0.0, 0.0);
    double4 f = OMG ? zero4 : S.frc[a], fpa = OMG ? zero4 : S.fp[a], fva = OMG ? zero4 : S.fv[a];
    const int n = nbcnt[a];
    constexpr int U = 4;
    for (int k0 = 0; k0 < n; k0 += U) {
      SPHDEM_FETCH_ENTRIES(U)
      double4 p[U], v[U];
      double2 d[U], hi[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int j = e[u];
        p[u] = S.posh[j];
        v[u] = S.vhm[j];
        d[u] = *reinterpret_cast<const double2*>(&S.dpo[j]);  // dens, pdr2
        hi[u] = hinv[j];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (k0 + u >= n) break;
        const double hb = p[u].w, mb = v[u].w, rhob = d[u].x, pdr2b = d[u].y;
        const double dx = pa.x - p[u].x, dy = pa.y - p[u].y, dz = pa.z - p[u].z;
        const double l2 = dot3(dx, dy, dz, dx, dy, dz);
        const double r = sqrt(l2);
        const double rinv = r > 0.0 ? 1.0 / r : 0.0;  // r = 0: dr = 0, gradW = 0
        const double qa = divr(r, ha, rha), qb = divr(r, hb, hi[u].x);
        if (OMG) om += mb * wendPdWhi(P.wcon, qa, ha, rha, i3, c3);
        const double Kb = wrom * (hi[u].y * hb);  // WCON hb^-4 (/ omega_a)
        double Ga = qa <= 2.0 ? wendPoly(qa) * Ka * rinv : 0.0;
        double Gb = qb <= 2.0 ? wendPoly(qb) * Kb * rinv : 0.0;
        if (!VARH) Ga = Gb = 0.5 * (Ga + Gb);  // the average gradient
        if (fluid) {
          const double dv0 = va.x - v[u].x, dv1 = va.y - v[u].y, dv2 = va.z - v[u].z;
          const double Gbar = 0.5 * (Ga + Gb);  // gradW-bar = dr * Gbar
          const double hav = 0.5 * (ha + hb);
          double fvs = 0.0;  // fv = dv * fvs (Morris)
          double fvr = 0.0;  // fv = dr * fvr (Cleary, Monaghan: m_b visc gradW-bar)
          double vsig = vsigC;
          const double vdr = (ARTIF || VISC == VISC_CLE || VISC == VISC_MON)
                                 ? dot3(dx, dy, dz, dv0, dv1, dv2)
                                 : 0.0;
          if (VISC == VISC_MOR) {
            const double rab = 1.0 / (da.x * rhob);
            const double visc = LIQDEM ? viscLiq * rab : P.viscosity * (da.x + rhob) * rab;
            const double den = l2 + 0.01 * (hav * hav);
            fvs = mb * visc * (Gbar * l2) / den;  // dr . gradW-bar = Gbar |dr|^2
          }
          if (VISC == VISC_CLE) {  // sph.h:598-604
            const double den = l2 + 0.01 * (hav * hav);
            const double visc = 19.8 * P.viscosity * vdr / ((da.x + rhob) * den);
            fvr = mb * visc * Gbar;
          }
          if (VISC == VISC_MON) {  // sph.h:606-630, no SLK
            const double viss = vdr * rinv;
            if (!LIQDEM) vsig = 2.0 * P.spsound + 2.0 * fabs(viss);
            const double visc = viss * vsig * alphaM / (da.x + rhob);
            fvr = mb * visc * Gbar;
          }
          if (ARTIF) {
            double avsig = 0, avisc = 0.0;
            if (vdr < 0) {
              const double viss = vdr * rinv;
              avsig = 2.0 * P.spsound - 4.0 * viss;
              avisc = viss * avsig * P.alpha_artificial / (da.x + rhob);
            }
            const double sg = mb * avisc * Gbar;
            f.x += sg * dx;
            f.y += sg * dy;
            f.z += sg * dz;
            vsig = dmax2(vsig, avsig);
            newDt = dmin2(newDt, 0.6 * hav / vsig);
          } else if (VISC == VISC_MON) {
            newDt = dmin2(newDt, 0.6 * hav / vsig);
          } else {
            minA = dmin2(minA, 0.6 * hav);
          }
          minB = dmin2(minB, cdiff * (hav * hav));
          constexpr bool alongDr = VISC == VISC_CLE || VISC == VISC_MON;
          const double fv0 = alongDr ? dx * fvr : dv0 * fvs;
          const double fv1 = alongDr ? dy * fvr : dv1 * fvs;
          const double fv2 = alongDr ? dz * fvr : dv2 * fvs;
          fva.x += fv0;
          fva.y += fv1;
          fva.z += fv2;
          f.x += fv0;
          f.y += fv1;
          f.z += fv2;
        }
        // calcPressForce (sph.h:698-710), antic = 0 under NO_ANTICLUMPING:
        // -mb (pdr2a gradWa + pdr2b gradWb) = dr * pp
        const double pp = -mb * (da.y * Ga + pdr2b * Gb);
        const double p0 = pp * dx, p1 = pp * dy, p2 = pp * dz;
        fpa.x += p0;
        fpa.y += p1;
        fpa.z += p2;
        f.x += p0;
        f.y += p1;
        f.z += p2;
      }
    }
    if (!ARTIF && VISC != VISC_MON && minA != inf) newDt = dmin2(newDt, minA / vsigC);
    if (minB != inf) newDt = dmin2(newDt, minB / P.viscosity);
    if (OMG) {
      const double omega = 1 + ha * om / (da.x * 3);
      S.dpo[a].z = omega;
      const double rw = 1.0 / omega;
      const double4 f0 = S.frc[a], fp0 = S.fp[a], fv0 = S.fv[a];
      f = make_double4(f0.x + rw * f.x, f0.y + rw * f.y, f0.z + rw * f.z, f0.w);
      fpa = make_double4(fp0.x + rw * fpa.x, fp0.y + rw * fpa.y, fp0.z + rw * fpa.z, fp0.w);
      fva = make_double4(fv0.x + rw * fva.x, fv0.y + rw * fva.y, fv0.z + rw * fva.z, fv0.w);
    }
    S.frc[a] = f;
    S.fp[a] = fpa;
    S.fv[a] = fva;
  }
  atomicMinPos(&g->newDt_bits, newDt);
}

}  // namespace sphdem
