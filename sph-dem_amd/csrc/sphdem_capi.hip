// sphdem_capi.hip -- C ABI of the gfx950 SPH-DEM loop (include/sphdem_c.h):
// device context, cell build orchestration (CdataLL::reset), the three
// half-steps of CsphIncompress and the parity entry points.  All work is
// ordered on the context's stream; the only host synchronisation inside a step
// is the read-back of hmax in reset(), which sizes the cell grid.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <utility>
#include <atomic>
#include <chrono>
#include <mutex>
#include <cmath>
#include <thread>
#include <vector>

#include "sphdem_kernels.hip"
#include "sphdem_tiles.hip"
#include "sphdem_mixed.hip"
#include "sphdem_adjoint.hip"

using namespace sphdem;

// The host-side fields a step changes: kept around a capture (which runs a
// step's host code without running the step) and set after its replays.
struct HostFields {
  long long sphStep;
  int ng, ns, nd, K, lgen, countsClean, gcntZeroAt;
  unsigned gstale;
  bool listsValid, segsValid, gridValid, recFresh, pvFresh, dFresh;
  double ssr;
  Grid gs, gc;
  GhostBand band;
  Store cur, alt;
  void* storeMem[2];
};

struct sphdem_ctx {
  sphdem_params P;
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream for passes that run beside the main chain, and its events
  hipStream_t side = nullptr;
  hipEvent_t evFork = nullptr, evJoin = nullptr;
  bool overlap = true;  // run the independent passes on the side stream
  int64_t cap = 0;
  int n = 0, ns0 = 0, nd0 = 0, ns = 0, nd = 0;
  Store cur{}, alt{};
  void* storeMem[2] = {nullptr, nullptr};
  int *key = nullptr, *slot = nullptr, *tmp = nullptr, *perm = nullptr;
  int *counts = nullptr, *start = nullptr, *tiles = nullptr;
  int Kcap = 0, tileCap = 0;
  // the pair export's per-query counts and offsets (grown on demand)
  int* pairCnt = nullptr;
  long long* pairOff = nullptr;
  int pairCap = 0;
  // segments of the SPH grid (the wavefront work units of the pair passes)
  int *segCnt = nullptr, *segOff = nullptr;
  Segment* segs = nullptr;
  int segColCap = 0, segCap = 0;
  bool segsValid = false;  // the segments partition the current store
  DevGlobals* hostG = nullptr;  // pinned host copy of dg for the reset's readback
  // tile shape: particles per list-build segment (<= 64, one wavefront; 0 for
  // the per-target search) and threads per workgroup of the SPH pair walkers
  // (sphdem_set_tile_shape)
  int tileSeg = 0, walkTPB = 256;
  // cached SPH neighbour lists (valid from their first use until the next reset)
  int *nbl = nullptr, *nbcnt = nullptr;
  // the per-target search's quantised records (one per entry, plus slack for
  // a batch read past a range's end) and its two generation flag slots
  uint2* lrec = nullptr;
  int* lflag = nullptr;
  int lgen = 0;
  // DEM cells with a sphere within one cell, marked with the number in nearGen
  int* demNear = nullptr;
  int demNearCap = 0;
  bool listsValid = false;
  // the list search's records (recFresh) and the mixed form's P and V records
  // (pvFresh) were written by the last reset's permutation and no position,
  // h or vhat changed since
  bool recFresh = false, pvFresh = false;
  // the mixed form's D records were written by the last EOS pass and no
  // density, pdr2 or h changed since (a box without ghosts)
  bool dFresh = false;
  double2* hinv = nullptr;  // {1/h, 1/h^5} per SPH-group particle
  long long* okey = nullptr;  // order key of each entry, in scatter order
  // mixed form: packed neighbour records per entry (allocated on first use)
  uint4* mixP = nullptr;
  float4 *mixV = nullptr, *mixD = nullptr;
  int64_t mixCap = 0;
  float mixRes = 0.0f;  // fixed-point unit of the last packing
  double4* posm = nullptr;  // FP64 {x, y, z, mass} per entry for omega/density
  int64_t posmCap = 0;
  // periodic ghosts (store entries [n, n + ng) after a reset)
  int ng = 0, gbase = 0;    // ghost count; first key of the ghost cells (0: none)
  // ghost field groups whose origins changed since the ghosts were copied
  unsigned gstale = 7;
  int gcntZeroAt = -1;  // the n for which gcnt[n] was cleared
  int countsClean = 0;  // counts[0, countsClean) are zero (cleared by k_scatter)
  GhostBand band{};         // cells a ghost may occupy on the current grid
  bool gridValid = false;   // gs/band describe the grid of the last reset
  int64_t ghostCap = 0;     // store entries beyond cap available to ghosts
  int *gcnt = nullptr, *goff = nullptr;                   // per real entry (n + 1)
  int *gsrc = nullptr, *gord = nullptr, *gcell = nullptr;  // per ghost, before the permutation
  int *gorig = nullptr, *gcode = nullptr;                  // per ghost, after it
  int* inv = nullptr;                                      // old -> new position of real entries
  double* partial = nullptr;
  int partialCap = 0;
  double* fieldbuf = nullptr;
  int* errflag = nullptr;
  DevGlobals* dg = nullptr;
  long long sphStep = 0;
  double ssr = 0.0;
  Grid gs{}, gc{};
  Box box{};
  CouplingRule rule{};
  int K = 0, reachShep = 1, reachFull = 1;
  bool loaded = false;
  // measurement window: events bracketing each launch of one kernel
  int timedKernel = -1;
  std::vector<hipEvent_t> evPool;
  size_t evUsed = 0;
  char err[512] = {0};
  // steps replayed from a captured graph (sphdem_run_steps_graph)
  bool graphMode = false;  // capturing: the resets read nothing back
  int graphG = 0;          // ghost entries a captured reset processes
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  std::vector<unsigned char> graphKey;  // what the captured launches depend on
  long long graphSerial = 0;            // which capture graph holds
  int* gate = nullptr;      // device: stop flag, the replay that set it, replays started
  int* gateHost = nullptr;  // pinned: the gate and the live ghost count
  Store keep{};             // the state a replayed step started from
  void* keepMem = nullptr;
  DevGlobals* dgKeep = nullptr;
  int* nearGen = nullptr;     // device: the number of the DEM-cell marks
  long long allocEpoch = 0;   // counts buffer (re)allocations
  long long readyEpoch = -1;  // allocEpoch when a step as launches last ended
  long long deferredStep0 = 0;  // sphStep when the deferred steps began
  HostFields graphAfter{};    // the host fields after the captured step
};

namespace {
enum KernelId {
  KID_FORCE = 0, KID_OMEGA, KID_DDDT, KID_COUPLING, KID_SHEPSUM, KID_FPI, KID_CONTACTS,
  KID_EOS, KID_START, KID_KICK, KID_REST, KID_BIN, KID_RANK, KID_PERMUTE, KID_WRAP, KID_LISTS,
  KID_COUNT
};
const char* kKernelNames[KID_COUNT] = {"force", "omega", "dddt", "coupling", "shepsum",
                                       "fpi",   "contacts", "eos", "start", "kick",
                                       "rest",  "bin",   "rank",  "permute", "wrap",
                                       "lists"};
constexpr size_t kMaxTimedEvents = 40000;

bool recordEvent(sphdem_ctx* c) {
  if (c->evUsed == c->evPool.size()) {
    if (c->evPool.size() >= kMaxTimedEvents) return false;
    hipEvent_t e;
    // timestamps only (nothing synchronizes with these events but the final
    // elapsed-time readout after a stream synchronize): no system-scope
    // acquire/release, whose cache writeback and invalidation would cost the
    // stream more than the kernels timed
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return false;
    c->evPool.push_back(e);
  }
  return hipEventRecord(c->evPool[c->evUsed++], c->stream) == hipSuccess;
}
// RAII bracket of one kernel launch (records only inside an open window)
struct Timed {
  sphdem_ctx* c;
  bool on;
  Timed(sphdem_ctx* c_, int kid) : c(c_), on(c_->timedKernel == kid && c_->evUsed % 2 == 0) {
    if (on) on = recordEvent(c);
  }
  ~Timed() {
    if (on && !recordEvent(c)) c->evUsed--;  // keep start/stop events paired
  }
};
}  // namespace

namespace {

int fail(sphdem_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIPC(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, SPHDEM_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
#define LAUNCHED()                                                                       \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess)                                                                \
      return fail(ctx, SPHDEM_E_HIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
  } while (0)
#define CHECK(expr)     \
  do {                  \
    int r_ = (expr);    \
    if (r_) return r_;  \
  } while (0)

constexpr int TPB = 256;
inline int blocks(int n) { return (n + TPB - 1) / TPB; }
// workgroups of an SPH pair walker over n particles (the tile shape's block)
inline int walkBlocks(const sphdem_ctx* ctx, int n) {
  return (n + ctx->walkTPB - 1) / ctx->walkTPB;
}
// grid of a grid-stride kernel that ends in one atomic per block on a single
// address: those atomics serialise, so at most two blocks per CU
inline int reduceBlocks(int n) { return std::min(blocks(n), 512); }
// kernels that give each item a whole 64-lane wavefront
inline int waveBlocks(int n) { return (n + TPB / 64 - 1) / (TPB / 64); }

// The store keeps kStoreSlack entries after its capacity: the list build reads
// candidates in whole batches of four past the end of a range.
constexpr int kStoreSlack = 4;
int allocStore(sphdem_ctx* ctx, Store& s, void*& mem, int64_t cap) {
  cap += kStoreSlack;
  const size_t vec = sizeof(double4) * (size_t)cap;
  const size_t total = vec * kNumVec + 3 * sizeof(int) * (size_t)cap;
  HIPC(hipMalloc(&mem, total));
  char* p = (char*)mem;
  double4** f[kNumVec] = {&s.posh, &s.vhm, &s.dpo, &s.vel, &s.v0p, &s.vh0d,
                          &s.frc,  &s.fp,  &s.fv,  &s.fb,  &s.fdrag, &s.fvel};
  for (int k = 0; k < kNumVec; k++) {
    *f[k] = (double4*)p;
    p += vec;
  }
  s.iam = (int*)p;
  p += sizeof(int) * (size_t)cap;
  s.orig = (int*)p;
  p += sizeof(int) * (size_t)cap;
  s.cell = (int*)p;
  return 0;
}

double4* storeArr(Store& s, int k) {
  double4* a[kNumVec] = {s.posh, s.vhm, s.dpo, s.vel, s.v0p, s.vh0d,
                         s.frc,  s.fp,  s.fv,  s.fb,  s.fdrag, s.fvel};
  return a[k];
}

// Per-entry buffers (the two stores, the sort scratch, hinv) for cap real
// entries plus ghostCap ghosts; the ghost records sized ghostCap.  On growth
// the first n entries of the current store and of key/slot are kept.
int allocEntries(sphdem_ctx* ctx, int64_t ghostCap) {
  const int64_t ent = ctx->cap + ghostCap;
  Store cur{}, alt{};
  void* mem[2] = {nullptr, nullptr};
  CHECK(allocStore(ctx, cur, mem[0], ent));
  CHECK(allocStore(ctx, alt, mem[1], ent));
  int *key = nullptr, *slot = nullptr, *tmp = nullptr, *perm = nullptr;
  double2* hinv = nullptr;
  long long* okey = nullptr;
  HIPC(hipMalloc(&okey, sizeof(long long) * ent));
  HIPC(hipMalloc(&key, sizeof(int) * ent));
  HIPC(hipMalloc(&slot, sizeof(int) * ent));
  HIPC(hipMalloc(&tmp, sizeof(int) * ent));
  HIPC(hipMalloc(&perm, sizeof(int) * ent));
  HIPC(hipMalloc(&hinv, sizeof(double2) * ent));
  uint2* lrec = nullptr;
  HIPC(hipMalloc(&lrec, sizeof(uint2) * (ent + 64)));
  if (ctx->storeMem[0]) {
    const size_t n = (size_t)ctx->n;
    for (int k = 0; k < kNumVec; k++)
      HIPC(hipMemcpyAsync(storeArr(cur, k), storeArr(ctx->cur, k), sizeof(double4) * n,
                          hipMemcpyDeviceToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(cur.iam, ctx->cur.iam, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(cur.orig, ctx->cur.orig, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(cur.cell, ctx->cur.cell, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(key, ctx->key, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(slot, ctx->slot, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
  }
  void* old[] = {ctx->storeMem[0], ctx->storeMem[1], ctx->key, ctx->slot, ctx->tmp, ctx->perm,
                 ctx->hinv, ctx->okey, ctx->gsrc, ctx->gord, ctx->gcell, ctx->gorig,
                 ctx->gcode, ctx->lrec};
  for (void* q : old)
    if (q) HIPC(hipFree(q));
  ctx->cur = cur;
  ctx->alt = alt;
  ctx->storeMem[0] = mem[0];
  ctx->storeMem[1] = mem[1];
  ctx->key = key;
  ctx->slot = slot;
  ctx->tmp = tmp;
  ctx->perm = perm;
  ctx->hinv = hinv;
  ctx->okey = okey;
  ctx->lrec = lrec;
  ctx->ghostCap = ghostCap;
  ctx->allocEpoch++;
  HIPC(hipMalloc(&ctx->gsrc, sizeof(int) * ghostCap));
  HIPC(hipMalloc(&ctx->gord, sizeof(int) * ghostCap));
  HIPC(hipMalloc(&ctx->gcell, sizeof(int) * ghostCap));
  HIPC(hipMalloc(&ctx->gorig, sizeof(int) * ghostCap));
  HIPC(hipMalloc(&ctx->gcode, sizeof(int) * ghostCap));
  // valid entries everywhere: a captured step's passes read up to its bound
  HIPC(hipMemsetAsync(ctx->gorig, 0, sizeof(int) * ghostCap, ctx->stream));
  HIPC(hipMemsetAsync(ctx->gcode, 0, sizeof(int) * ghostCap, ctx->stream));
  return 0;
}

// where each field lives: array index in the Store, component offset, width
struct FieldLoc {
  int arr, off, width;
};
bool fieldLoc(int field, FieldLoc* L) {
  switch (field) {
    case SPHDEM_FIELD_R: *L = {0, 0, 3}; return true;
    case SPHDEM_FIELD_H: *L = {0, 3, 1}; return true;
    case SPHDEM_FIELD_VHAT: *L = {1, 0, 3}; return true;
    case SPHDEM_FIELD_MASS: *L = {1, 3, 1}; return true;
    case SPHDEM_FIELD_DENS: *L = {2, 0, 1}; return true;
    case SPHDEM_FIELD_PDR2: *L = {2, 1, 1}; return true;
    case SPHDEM_FIELD_OMEGA: *L = {2, 2, 1}; return true;
    case SPHDEM_FIELD_POROSITY: *L = {2, 3, 1}; return true;
    case SPHDEM_FIELD_V: *L = {3, 0, 3}; return true;
    case SPHDEM_FIELD_U: *L = {3, 3, 1}; return true;
    case SPHDEM_FIELD_V0: *L = {4, 0, 3}; return true;
    case SPHDEM_FIELD_PRESS: *L = {4, 3, 1}; return true;
    case SPHDEM_FIELD_VHAT0: *L = {5, 0, 3}; return true;
    case SPHDEM_FIELD_DDDT: *L = {5, 3, 1}; return true;
    case SPHDEM_FIELD_F: *L = {6, 0, 3}; return true;
    case SPHDEM_FIELD_DUDT: *L = {6, 3, 1}; return true;
    case SPHDEM_FIELD_FP: *L = {7, 0, 3}; return true;
    case SPHDEM_FIELD_SHEPSUM: *L = {7, 3, 1}; return true;
    case SPHDEM_FIELD_FV: *L = {8, 0, 3}; return true;
    case SPHDEM_FIELD_FB: *L = {9, 0, 3}; return true;
    case SPHDEM_FIELD_FDRAG: *L = {10, 0, 3}; return true;
    case SPHDEM_FIELD_FVEL: *L = {11, 0, 3}; return true;
    default: return false;
  }
}

bool liqDem(const sphdem_ctx* c) { return (c->P.flags & SPHDEM_F_LIQ_DEM) != 0; }
bool varH(const sphdem_ctx* c) { return (c->P.flags & SPHDEM_F_VAR_H_CORRECTION) != 0; }

int ensureKeys(sphdem_ctx* ctx, int K) {
  if (K + 1 > ctx->Kcap) {
    if (ctx->counts) HIPC(hipFree(ctx->counts));
    if (ctx->start) HIPC(hipFree(ctx->start));
    ctx->Kcap = (K + 1) + (K + 1) / 4;
    ctx->countsClean = 0;
    HIPC(hipMalloc(&ctx->counts, sizeof(int) * ctx->Kcap));
    HIPC(hipMalloc(&ctx->start, sizeof(int) * ctx->Kcap));
    ctx->allocEpoch++;
  }
  return 0;
}

// exclusive scan of in[0..m) into out (m <= SCAN_TILE^2); copyLast (device,
// optional) also receives out[m - 1]
int scanInts(sphdem_ctx* ctx, const int* in, int* out, int m, int* copyLast = nullptr) {
  const int tiles = (m + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles > SCAN_TILE) return fail(ctx, SPHDEM_E_CAPACITY, "scan of %d items", m);
  if (tiles + 1 > ctx->tileCap) {
    if (ctx->tiles) HIPC(hipFree(ctx->tiles));
    ctx->tileCap = tiles + 1 + (tiles + 1) / 4;
    HIPC(hipMalloc(&ctx->tiles, sizeof(int) * ctx->tileCap));
    ctx->allocEpoch++;
  }
  hipLaunchKernelGGL(k_scan_tile, dim3(tiles), dim3(SCAN_THREADS), 0, ctx->stream, in, out,
                     ctx->tiles, m, tiles == 1 ? copyLast : (int*)nullptr);
  LAUNCHED();
  if (tiles > 1) {
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(SCAN_THREADS), 0, ctx->stream, out,
                       ctx->tiles, m, copyLast);
    LAUNCHED();
  }
  return 0;
}
int scanCounts(sphdem_ctx* ctx, int m) { return scanInts(ctx, ctx->counts, ctx->start, m); }

// Segments of the SPH grid after a rebuild: per-column counts, a scan, then
// the segment records in column order (so segment order follows space).
int buildSegments(sphdem_ctx* ctx) {
  ctx->segsValid = ctx->tileSeg != 0;
  if (ctx->tileSeg == 0) return 0;  // the per-target search needs no segments
  const Grid& g = ctx->gs;
  const int ncol = g.dim[0] * g.dim[1];
  if (ncol + 1 > ctx->segColCap) {
    if (ctx->segCnt) HIPC(hipFree(ctx->segCnt));
    if (ctx->segOff) HIPC(hipFree(ctx->segOff));
    ctx->segColCap = (ncol + 1) + (ncol + 1) / 4;
    HIPC(hipMalloc(&ctx->segCnt, sizeof(int) * ctx->segColCap));
    HIPC(hipMalloc(&ctx->segOff, sizeof(int) * ctx->segColCap));
    ctx->allocEpoch++;
  }
  // a column holds at most ceil(n_col/64) + (nonempty cells) segments
  const int need = ctx->ns + ncol + 1;
  if (need > ctx->segCap) {
    if (ctx->segs) HIPC(hipFree(ctx->segs));
    ctx->segCap = need + need / 4;
    HIPC(hipMalloc(&ctx->segs, sizeof(Segment) * ctx->segCap));
    ctx->allocEpoch++;
  }
  // one wavefront per column; the count kernel's waves cover ncol + 1 (the
  // last one writes the scan's closing zero)
  hipLaunchKernelGGL(k_seg_count, dim3((ncol + 1 + kSegWaves - 1) / kSegWaves),
                     dim3(64 * kSegWaves), 0, ctx->stream, ctx->start, g, ctx->tileSeg,
                     ctx->segCnt);
  LAUNCHED();
  CHECK(scanInts(ctx, ctx->segCnt, ctx->segOff, ncol + 1));
  hipLaunchKernelGGL(k_seg_write, dim3((ncol + kSegWaves - 1) / kSegWaves), dim3(64 * kSegWaves),
                     0, ctx->stream, ctx->start, g, ctx->tileSeg, ctx->segOff, ctx->segs);
  LAUNCHED();
  return 0;
}
// workgroups of the tile kernels: a multiple of 8 (one slab per XCD)
int tileBlocks(const sphdem_ctx* ctx) {
  int est = (ctx->ns / 32 + kWavesPerBlock - 1) / kWavesPerBlock;
  est = (est + 7) & ~7;
  return std::min(std::max(est, 8), 2048);
}
// Ghost copies of the fields the pair passes read of neighbours, in three
// groups: posh (r, h), vhm (vhat, mass), dpo (dens, pdr2, omega, porosity).
// A launch that may write a group of real SPH-group particles marks it stale;
// a pass refreshes the ghosts when a group it reads of neighbours is stale.
constexpr unsigned GS_POSH = 1, GS_VHM = 2, GS_DPO = 4, GS_ALL = 7;
void markStale(sphdem_ctx* ctx, unsigned groups) {
  ctx->gstale |= groups;
  if (groups & (GS_POSH | GS_VHM)) ctx->recFresh = ctx->pvFresh = false;
  if (groups & (GS_POSH | GS_DPO)) ctx->dFresh = false;
}
int refreshGhosts(sphdem_ctx* ctx, unsigned need) {
  if (ctx->ng == 0 || !(ctx->gstale & need)) return 0;
  hipLaunchKernelGGL(k_ghost_refresh, dim3(blocks(ctx->ng)), dim3(TPB), 0, ctx->stream, ctx->cur,
                     ctx->n, ctx->ng, ctx->gorig, ctx->gcode, ctx->box,
                     ctx->graphMode ? &ctx->dg->ghostTotal : (const int*)nullptr);
  LAUNCHED();
  ctx->gstale = 0;
  return 0;
}
// The quantisation unit of the per-target search: a power of two u with
// edge / u in [1024, 2048), origin the grid's lower corner.
ListQuant listQuant(const Grid& g) {
  int e = 0;
  (void)frexp(g.edge, &e);  // edge = m 2^e, m in [0.5, 1)
  ListQuant q;
  for (int d = 0; d < 3; d++) q.o[d] = g.cmin[d];
  q.uinv = ldexp(1.0, 11 - e);
  return q;
}
int launchLists(sphdem_ctx* ctx) {
  CHECK(refreshGhosts(ctx, GS_POSH));
  if (ctx->tileSeg == 0) {
    if (ctx->ns == 0) return 0;
    const ListQuant lq = listQuant(ctx->gs);
    const int gen = ctx->lgen++ & 1;
    if (!ctx->recFresh) {  // else the reset's permutation wrote them
      hipLaunchKernelGGL(k_lrec, dim3(blocks(ctx->ns + ctx->ng)), dim3(TPB), 0, ctx->stream,
                         ctx->cur.posh, ctx->ns, ctx->n, ctx->ng, lq, ctx->lrec, ctx->lflag, gen,
                         ctx->graphMode ? &ctx->dg->ghostTotal : (const int*)nullptr);
      LAUNCHED();
    }
    ctx->recFresh = false;
    const int per = (ctx->ns + 7) / 8;
    hipLaunchKernelGGL(k_lists_t<32>, dim3(8 * ((per + TPB - 1) / TPB)), dim3(TPB), 0,
                       ctx->stream, ctx->cur.posh, ctx->lrec, ctx->cur.cell, ctx->gs, ctx->gbase,
                       ctx->start, ctx->ns, lq.uinv, ctx->lflag, gen, ctx->nbl, ctx->nbcnt,
                       ctx->dg);
    LAUNCHED();
    return 0;
  }
  if (!ctx->segsValid) CHECK(buildSegments(ctx));
  hipLaunchKernelGGL(k_lists, dim3(tileBlocks(ctx)), dim3(TPB), 0, ctx->stream, ctx->cur.posh,
                     ctx->cur.cell, ctx->gs, ctx->gbase, ctx->start, ctx->segs,
                     ctx->segOff + ctx->gs.dim[0] * ctx->gs.dim[1], ctx->ns, ctx->nbl,
                     ctx->nbcnt, ctx->dg);
  LAUNCHED();
  return 0;
}

void setupGrids(sphdem_ctx* ctx) {
  const sphdem_params& P = ctx->P;
  Grid& gs = ctx->gs;
  for (int d = 0; d < 3; d++) gs.cmin[d] = P.proc_domain[2 * d];
  // getCellI (dataLL.h:303-305) and the cell array size (dataLL.cpp:16, 222)
  gs.edge = 2.0 * ctx->ssr;
  gs.off = 3.0;
  for (int d = 0; d < 3; d++)
    gs.dim[d] = (int)((P.proc_domain[2 * d + 1] - P.proc_domain[2 * d]) / gs.edge) + 8;
  gs.base = 0;
  const int Cs = gs.dim[0] * gs.dim[1] * gs.dim[2];
  Grid& gc = ctx->gc;
  for (int d = 0; d < 3; d++) gc.cmin[d] = P.proc_domain[2 * d];
  if (liqDem(ctx)) {
    gc.edge = std::max(P.coupling_radius, 2.0 * P.dem_search_radius) * (1.0 + 1e-9);
    gc.off = 1.0;
    for (int d = 0; d < 3; d++)
      gc.dim[d] = (int)((P.proc_domain[2 * d + 1] - P.proc_domain[2 * d]) / gc.edge) + 3;
  } else {
    gc.edge = 1.0;
    gc.off = 0.0;
    gc.dim[0] = gc.dim[1] = gc.dim[2] = 1;
  }
  gc.base = Cs;
  gs.inv = 1.0 / gs.edge;  // the kernels' cell indices (cellTrunc)
  gc.inv = 1.0 / gc.edge;
  const int Cc = gc.dim[0] * gc.dim[1] * gc.dim[2];
  // keys: SPH cells, DEM cells, one for deleted particles, then (periodic
  // boxes) the SPH cells again for the ghosts
  const bool periodic = ctx->box.periodic[0] || ctx->box.periodic[1] || ctx->box.periodic[2];
  ctx->gbase = periodic ? Cs + Cc + 1 : 0;
  ctx->K = periodic ? Cs + Cc + 1 + Cs : Cs + Cc + 1;
  if (liqDem(ctx)) {
    const double Rc = P.coupling_radius;
    ctx->reachShep = std::max(1, (int)ceil(Rc * (1.0 + 1e-9) / gs.edge));
    ctx->reachFull =
        (int)ceil((ctx->rule.nc + 1) * ctx->rule.e / gc.edge) + 1;
  }
}

void setupStatic(sphdem_ctx* ctx) {
  const sphdem_params& P = ctx->P;
  for (int d = 0; d < 3; d++) {
    ctx->box.periodic[d] = P.periodic[d] ? 1 : 0;
    ctx->box.Lp[d] = P.rmax[d] - P.rmin[d];
    ctx->box.Lm[d] = P.rmin[d] - P.rmax[d];
  }
  CouplingRule& R = ctx->rule;
  if (P.dem_search_radius > 0 && P.coupling_radius > 0) {
    const double Rc = P.coupling_radius, DSR = P.dem_search_radius;
    for (int d = 0; d < 3; d++) R.cmin[d] = P.proc_domain[2 * d];
    R.e = 2 * DSR;
    R.einv = 1.0 / R.e;
    R.s = 2 * Rc / DSR;
    const double t = 0.5 * Rc / DSR;
    R.thr = t * t;
    R.rc2 = Rc * Rc;
    R.skip2 = R.rc2 * (1.0 + 1e-12);
    R.nc = (int)ceil(0.5 * Rc / DSR);
    for (int d = 0; d < 3; d++) {
      const double dmin = P.proc_domain[2 * d];
      const double dspace = P.proc_domain[2 * d + 1] - dmin;
      R.gmin[d] = dmin + Rc;
      R.gspace[d] = (dspace - Rc) - Rc;
    }
  }
}

// CdataLL::reset (dataLL.cpp:140-236)
// the ghost band of the current grid: the cells of the domain's extreme
// coordinates (every wrapped particle lies between them), widened by one
void setupBand(sphdem_ctx* ctx) {
  const Grid& g = ctx->gs;
  for (int d = 0; d < 3; d++) {
    const double lo = ctx->P.proc_domain[2 * d], hi = ctx->P.proc_domain[2 * d + 1];
    const int clo = (int)((lo - g.cmin[d]) / g.edge + g.off);
    const int chi = (int)((hi - g.cmin[d]) / g.edge + g.off);
    ctx->band.lo[d] = std::max(clo - 1, 0);
    ctx->band.hi[d] = std::min(chi + 1, g.dim[d] - 1);
  }
}

// keepSsr: sort only, keeping the search radius (a checkpoint restore, which
// must not move the hysteresis of calcDomainLimits)
// ---- work beside the main chain
// Side(ctx) makes the side stream wait for everything issued so far on the
// main stream and routes the context's launches to the side stream until
// back() (or the destructor); the main stream then waits, at join() (or the
// destructor), for everything the side stream was given.  Every entry point
// therefore returns with the side work ordered before the main stream's
// next operation.
struct Side {
  sphdem_ctx* c;
  hipStream_t mainStream;
  bool onSide = false, joined = false;
  // create: make the side stream if there is none yet; otherwise only an
  // existing one is used.  The side stream is made by the step passes on
  // first use, so a context that never overlaps (the boxes of an ensemble)
  // keeps one stream: HIP deals streams to its hardware queues in creation
  // order, and a second stream per box would halve the queues the boxes'
  // main streams get.
  explicit Side(sphdem_ctx* c_, bool create = true) : c(c_), mainStream(c_->stream) {
    if (c->overlap && !c->side && create &&
        (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming | hipEventReleaseToDevice) !=
             hipSuccess ||
         hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming | hipEventReleaseToDevice) !=
             hipSuccess))
      c->overlap = false;
    if (c->overlap && c->side && hipEventRecord(c->evFork, mainStream) == hipSuccess &&
        hipStreamWaitEvent(c->side, c->evFork, 0) == hipSuccess) {
      c->stream = c->side;
      onSide = true;
    } else {
      joined = true;  // no overlap: everything stays on the main stream
    }
  }
  void back() {
    if (onSide) c->stream = mainStream;
    onSide = false;
  }
  int join() {
    back();
    if (joined) return 0;
    joined = true;
    sphdem_ctx* ctx = c;
    HIPC(hipEventRecord(ctx->evJoin, ctx->side));
    HIPC(hipStreamWaitEvent(mainStream, ctx->evJoin, 0));
    return 0;
  }
  ~Side() { (void)join(); }
};

// Error bits the kernels raise in DevGlobals::errflags (bit 0: k_bin found a
// particle outside its cell grid; bit 1: k_lists clipped a list at kMaxNb
// entries).  Reported once, then cleared, so that a caller that restores a
// checkpoint can continue.
int flagError(sphdem_ctx* ctx, unsigned flags) {
  if (!flags) return 0;
  HIPC(hipMemsetAsync(&ctx->dg->errflags, 0, sizeof(unsigned), ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  if (flags & 1u)
    return fail(ctx, SPHDEM_E_CAPACITY,
                "a particle fell outside the cell grid; the pair sets since the last reset are "
                "incomplete");
  return fail(ctx, SPHDEM_E_CAPACITY,
              "a particle had more than %d SPH neighbours; the passes since the last list build "
              "used a clipped list",
              kMaxNb);
}
// Reads the error bits after everything issued so far (one host round trip):
// the end of every entry point that builds neighbour lists, so that no call
// returns 0 after a pass used a clipped list.
int checkFlags(sphdem_ctx* ctx) {
  HIPC(hipMemcpyAsync(&ctx->hostG->errflags, &ctx->dg->errflags, sizeof(unsigned),
                      hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return flagError(ctx, ctx->hostG->errflags);
}

bool mixed(const sphdem_ctx* c);
Quant mixQuant(const Grid& g);
// midInit: middle()'s reset, which also starts the dt minimum at maxdt
// fold: the half-step before the reset, done by k_wrap and k_permute
// (WrapFold); its permutation also writes the records the passes after a
// step's reset pack from the store (recFresh, pvFresh)
int doReset(sphdem_ctx* ctx, bool keepSsr = false, bool midInit = false,
            const WrapFold& fold = WrapFold{}) {
  const sphdem_params& P = ctx->P;
  const int n = ctx->n;
  // a captured reset reads nothing back: it needs the steady state of a
  // reset on the same grid (buffers allocated, counts cleared, the ghost
  // count kept on the device and bounded by graphG) and leaves the checks
  // the host would make to k_reset_gate
  const bool dev = ctx->graphMode;
  if (dev && (n == 0 || ctx->gcntZeroAt != n || !ctx->gridValid || ctx->countsClean < ctx->K + 1))
    return fail(ctx, SPHDEM_E_STATE, "a captured reset needs the state a reset on the same grid "
                                     "leaves");
  ctx->segsValid = false;
  ctx->recFresh = ctx->pvFresh = ctx->dFresh = false;
  // hmax_bits is cleared by the previous reset's k_bin (or by the upload);
  // gcnt[n], the scans' closing zero, is written by nothing else
  if (n == 0) HIPC(hipMemsetAsync(&ctx->dg->hmax_bits, 0, sizeof(unsigned long long), ctx->stream));
  if (ctx->gcntZeroAt != n) {
    HIPC(hipMemsetAsync(ctx->gcnt + n, 0, sizeof(int), ctx->stream));
    ctx->gcntZeroAt = n;
  }
  double dmin[3], dsp[3];
  for (int d = 0; d < 3; d++) {
    dmin[d] = P.proc_domain[2 * d];
    dsp[d] = P.proc_domain[2 * d + 1] - dmin[d];
  }
  // With a periodic box the ghosts are counted on the current grid while
  // wrapping; the count stands if the grid survives this reset (one host
  // round trip per reset), and is redone on the new grid otherwise.
  const bool periodic = ctx->box.periodic[0] || ctx->box.periodic[1] || ctx->box.periodic[2];
  const bool pre = periodic && ctx->gridValid && n > 0;
  int ngPre = 0;
  {
    // a folded half-step moves more bytes per particle: more waves in flight
    // (and at least one block, which advances the time)
    const int wb = fold.mode ? std::max(1, std::min(blocks(n), 4096)) : reduceBlocks(n);
    Timed tm(ctx, KID_WRAP);
    hipLaunchKernelGGL(k_wrap, dim3(wb), dim3(TPB), 0, ctx->stream, ctx->cur, n, dmin[0],
                       dmin[1], dmin[2], dsp[0], dsp[1], dsp[2], ctx->box, ctx->dg, ctx->gs,
                       ctx->band, pre ? ctx->gcnt : (int*)nullptr, midInit ? 1 : 0, fold);
    LAUNCHED();
  }
  if (pre) CHECK(scanInts(ctx, ctx->gcnt, ctx->goff, n + 1, &ctx->dg->ghostTotal));
  bool sameGrid = true;
  if (dev) {
    hipLaunchKernelGGL(k_reset_gate, dim3(1), dim3(1), 0, ctx->stream, ctx->dg, ctx->gate,
                       ctx->ssr, keepSsr ? 1 : 0,
                       (unsigned long long)(ctx->ns0 - ctx->ns),
                       (unsigned long long)(ctx->nd0 - ctx->nd), pre ? ctx->graphG : 0x7fffffff);
    LAUNCHED();
  } else {
    // into pinned memory: a direct copy, no staging through a runtime buffer
    HIPC(hipMemcpyAsync(ctx->hostG, ctx->dg, sizeof(DevGlobals), hipMemcpyDeviceToHost,
                        ctx->stream));
    const DevGlobals& hg = *ctx->hostG;
    HIPC(hipStreamSynchronize(ctx->stream));
    if (pre) ngPre = hg.ghostTotal;
    const double oldSsr = ctx->ssr;
    CHECK(flagError(ctx, hg.errflags));
    ctx->ns = ctx->ns0 - (int)hg.ndel_sph;
    ctx->nd = ctx->nd0 - (int)hg.ndel_dem;
    double hmax;
    unsigned long long hb = hg.hmax_bits;
    memcpy(&hmax, &hb, sizeof(hmax));
    if (ctx->ns + ctx->nd == 0) hmax = -100000;  // calcDomainLimits on an empty container
    if (!keepSsr && ((hmax > ctx->ssr) || (hmax < 0.9 * ctx->ssr))) ctx->ssr = hmax;
    setupGrids(ctx);
    setupBand(ctx);
    sameGrid = ctx->gridValid && ctx->ssr == oldSsr;
    ctx->gridValid = true;
  }
  for (const Grid* gr : {&ctx->gs, &ctx->gc})
    if (gr->dim[0] > 2048 || gr->dim[1] > 2048 || gr->dim[2] > 1024)
      return fail(ctx, SPHDEM_E_CAPACITY, "cell grid %d x %d x %d exceeds 2048 x 2048 x 1024",
                  gr->dim[0], gr->dim[1], gr->dim[2]);
  CHECK(ensureKeys(ctx, ctx->K));
  if (ctx->countsClean < ctx->K + 1)
    HIPC(hipMemsetAsync(ctx->counts, 0, sizeof(int) * (ctx->K + 1), ctx->stream));
  ctx->countsClean = 0;  // in use until k_scatter clears them
  const int keyDeleted = ctx->gs.dim[0] * ctx->gs.dim[1] * ctx->gs.dim[2] +
                         ctx->gc.dim[0] * ctx->gc.dim[1] * ctx->gc.dim[2];
  {
    Timed tm(ctx, KID_BIN);
    hipLaunchKernelGGL(k_bin, dim3(blocks(n)), dim3(TPB), 0, ctx->stream, ctx->cur, n, ctx->gs,
                       ctx->gc, keyDeleted, ctx->key, ctx->slot, ctx->counts, ctx->dg);
    LAUNCHED();
  }
  // ghosts (updateDomain's periodic images): count, size, then their keys
  int ng = 0;
  if (ctx->gbase > 0 && n > 0) {
    if (pre && sameGrid) {
      ng = dev ? ctx->graphG : ngPre;
    } else {
      hipLaunchKernelGGL(k_ghost_count, dim3(blocks(n)), dim3(TPB), 0, ctx->stream, ctx->cur, n,
                         ctx->gs, ctx->box, ctx->band, ctx->gcnt);
      LAUNCHED();
      CHECK(scanInts(ctx, ctx->gcnt, ctx->goff, n + 1));
      HIPC(hipMemcpyAsync(&ng, ctx->goff + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      HIPC(hipStreamSynchronize(ctx->stream));
    }
    if (ng > ctx->ghostCap) CHECK(allocEntries(ctx, (int64_t)ng + ng / 4 + 1024));
    if (ng > 0) {
      hipLaunchKernelGGL(k_ghost_make, dim3(blocks(n)), dim3(TPB), 0, ctx->stream, ctx->cur, n,
                         ctx->gs, ctx->box, ctx->band, ctx->goff, ctx->gbase, ctx->key, ctx->slot,
                         ctx->counts, ctx->gsrc, ctx->gord, ctx->gcell,
                         dev ? ctx->graphG : (int)ctx->ghostCap);
      LAUNCHED();
    }
  }
  const int ntot = n + ng;
  CHECK(scanCounts(ctx, ctx->K + 1));
  // the segments need only the cell starts: packed on the side stream (if
  // the steps have made one) while the main stream scatters, ranks and permutes
  Side side(ctx, false);
  CHECK(buildSegments(ctx));
  side.back();
  // a captured reset's ghost count is the device's (bounded by ng = graphG)
  const int* ngp = dev ? &ctx->dg->ghostTotal : nullptr;
  hipLaunchKernelGGL(k_scatter, dim3(blocks(ntot)), dim3(TPB), 0, ctx->stream, ntot, n, ctx->key,
                     ctx->slot, ctx->start, ctx->tmp, ctx->okey, ctx->cur.orig, ctx->gsrc,
                     ctx->gord, ctx->counts, ctx->K + 1, ngp);
  LAUNCHED();
  ctx->countsClean = ctx->K + 1;
  {
    Timed tm(ctx, KID_RANK);
    hipLaunchKernelGGL(k_rank, dim3(blocks(ntot)), dim3(TPB), 0, ctx->stream, ntot, n, ctx->key,
                       ctx->start, ctx->tmp, ctx->okey, ctx->perm, ngp);
    LAUNCHED();
  }
  PermRecords R{};
  if (fold.mode && ctx->ns > 0) {
    R.ns = ctx->ns;
    if (ctx->tileSeg == 0) {
      R.lists = 1;
      R.lq = listQuant(ctx->gs);
      R.rec = ctx->lrec;
      R.lflag = ctx->lflag;
      R.gen = ctx->lgen & 1;  // the generation the next list build takes
    }
    if (mixed(ctx) && ctx->mixCap >= ctx->cap + ctx->ghostCap) {
      R.mixed = 1;
      R.Q = mixQuant(ctx->gs);
      R.P = ctx->mixP;
      R.V = ctx->mixV;
    }
  }
  {
    Timed tm(ctx, KID_PERMUTE);
    if (fold.mode == 1)
      hipLaunchKernelGGL(k_permute<true>, dim3(blocks(ntot)), dim3(TPB), 0, ctx->stream, ntot, n,
                         ctx->perm, ctx->cur, ctx->alt, ctx->gsrc, ctx->gord, ctx->gcell,
                         ctx->box, ctx->inv, ngp, ctx->dg, fold, R);
    else
      hipLaunchKernelGGL(k_permute<false>, dim3(blocks(ntot)), dim3(TPB), 0, ctx->stream, ntot, n,
                         ctx->perm, ctx->cur, ctx->alt, ctx->gsrc, ctx->gord, ctx->gcell,
                         ctx->box, ctx->inv, ngp, ctx->dg, fold, R);
    LAUNCHED();
  }
  if (ng > 0) {
    hipLaunchKernelGGL(k_ghost_link, dim3(blocks(ng)), dim3(TPB), 0, ctx->stream, n, ng,
                       ctx->perm, ctx->inv, ctx->gsrc, ctx->gord, ctx->gorig, ctx->gcode, ngp);
    LAUNCHED();
  }
  ctx->ng = ng;
  ctx->gstale = 0;  // the permutation copied every ghost from its origin
  std::swap(ctx->cur, ctx->alt);
  std::swap(ctx->storeMem[0], ctx->storeMem[1]);
  ctx->listsValid = false;  // updateDomain clears every neighbour list (dataLL.cpp:374)
  ctx->recFresh = R.lists != 0;
  ctx->pvFresh = R.mixed != 0;
  return side.join();
}

// The cached lists are built by the first pair pass after a reset, with the h
// of that moment (dataLL.impl.h:186-188).
int ensureLists(sphdem_ctx* ctx) {
  if (ctx->listsValid) return 0;
  Timed tm(ctx, KID_LISTS);
  CHECK(launchLists(ctx));
  ctx->listsValid = true;
  return 0;
}

int checkLoaded(sphdem_ctx* ctx) {
  if (!ctx) return SPHDEM_E_ARG;
  if (!ctx->loaded) return fail(ctx, SPHDEM_E_STATE, "no particles uploaded");
  return 0;
}

// ---- passes
bool mixed(const sphdem_ctx* c) { return c->P.precision == SPHDEM_MIXED; }
// Packs {x, y, z, mass} of every SPH-group entry and ghost for the FP64
// omega and density sums.
int packPosm(sphdem_ctx* ctx) {
  const int64_t need = ctx->cap + ctx->ghostCap;
  if (ctx->posmCap < need) {
    if (ctx->posm) HIPC(hipFree(ctx->posm));
    ctx->posm = nullptr;
    ctx->posmCap = 0;
    HIPC(hipMalloc(&ctx->posm, sizeof(double4) * need));
    ctx->posmCap = need;
    ctx->allocEpoch++;
  }
  const int cnt = ctx->ns + ctx->ng;
  if (cnt > 0) {
    hipLaunchKernelGGL(k_pack_posm, dim3(blocks(cnt)), dim3(TPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->n, ctx->ng, ctx->posm);
    LAUNCHED();
  }
  return 0;
}
// The mixed form's fixed-point frame on grid g: the lower corner of cell 0
// and the power of two that spreads the grid's extent over 2^32 steps.
Quant mixQuant(const Grid& g) {
  Quant Q;
  double extent = 0.0;
  for (int d = 0; d < 3; d++) {
    Q.o[d] = g.cmin[d] - g.off * g.edge;  // lower corner of cell 0
    extent = std::max(extent, g.dim[d] * g.edge);
  }
  const int k = (int)std::ceil(std::log2(extent)) - 32;
  Q.inv = std::ldexp(1.0, -k);
  Q.res = std::ldexp(1.0f, k);
  return Q;
}
// Packs the mixed form's neighbour records (groups MIX_P/V/D) of every
// SPH-group entry and ghost, positions in fixed point over the current grid.
int packMix(sphdem_ctx* ctx, unsigned groups) {
  const int64_t need = ctx->cap + ctx->ghostCap;
  if (ctx->mixCap < need) {
    for (void* q : {(void*)ctx->mixP, (void*)ctx->mixV, (void*)ctx->mixD})
      if (q) HIPC(hipFree(q));
    ctx->mixP = nullptr;
    ctx->mixV = ctx->mixD = nullptr;
    ctx->mixCap = 0;
    HIPC(hipMalloc(&ctx->mixP, sizeof(uint4) * need));
    HIPC(hipMalloc(&ctx->mixV, sizeof(float4) * need));
    HIPC(hipMalloc(&ctx->mixD, sizeof(float4) * need));
    ctx->mixCap = need;
    ctx->pvFresh = ctx->dFresh = false;
    ctx->allocEpoch++;
  }
  const Quant Q = mixQuant(ctx->gs);
  ctx->mixRes = Q.res;
  if (ctx->pvFresh) groups &= ~(MIX_P | MIX_V);  // the reset's permutation wrote them
  if (ctx->dFresh) groups &= ~MIX_D;              // the EOS pass wrote them
  ctx->pvFresh = ctx->dFresh = false;
  const int cnt = ctx->ns + ctx->ng;
  if (cnt > 0 && groups) {
    hipLaunchKernelGGL(k_pack_mix, dim3(blocks(cnt)), dim3(TPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->n, ctx->ng, Q, groups, ctx->mixP, ctx->mixV, ctx->mixD);
    LAUNCHED();
  }
  return 0;
}
int passOmega(sphdem_ctx* ctx) {
  CHECK(ensureLists(ctx));
  CHECK(refreshGhosts(ctx, GS_POSH | GS_VHM));
  Timed tm(ctx, KID_OMEGA);
  if (mixed(ctx)) {
    CHECK(packMix(ctx, MIX_P));
    hipLaunchKernelGGL(k_omega_m, dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->nbl, ctx->nbcnt, ctx->mixP, ctx->mixRes, ctx->P.wcon);
  } else {
    CHECK(packPosm(ctx));
    hipLaunchKernelGGL(k_omega_l, dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->box, ctx->nbl, ctx->nbcnt, ctx->posm, ctx->P.wcon);
  }
  LAUNCHED();
  markStale(ctx, GS_DPO);
  return 0;
}
int passEos(sphdem_ctx* ctx) {
  Timed tm(ctx, KID_EOS);
  // the mixed D records too, when there are no ghosts (whose records the
  // packing takes after their refresh) and the record buffers exist
  float4* D = mixed(ctx) && ctx->ng == 0 && ctx->mixCap >= ctx->cap + ctx->ghostCap
                  ? ctx->mixD
                  : nullptr;
  hipLaunchKernelGGL(k_eos, dim3(blocks(ctx->ns)), dim3(TPB), 0, ctx->stream, ctx->cur, ctx->ns,
                     ctx->P, D);
  LAUNCHED();
  markStale(ctx, GS_DPO);
  ctx->dFresh = D != nullptr;
  return 0;
}
// {1/h, 1/h^5} of every SPH-group particle, for the passes that need them of
// neighbours (recomputed per pass: h changes in start, driftRest and the reinit)
int computeHinv(sphdem_ctx* ctx) {
  CHECK(refreshGhosts(ctx, GS_POSH));
  const int ntot = ctx->n + ctx->ng;  // ghosts included (their h is their origin's)
  hipLaunchKernelGGL(k_hinv, dim3(blocks(ntot)), dim3(TPB), 0, ctx->stream, ctx->cur.posh, ntot,
                     ctx->hinv);
  LAUNCHED();
  return 0;
}
// calcForce instantiated per flag combination (bit 0 VAR_H_CORRECTION, 1 VISC_ARTIFICIAL,
// 2 LIQ_DEM, bits 3-4 the viscosity variant); OMG: the walkers with calcOmega fused
// in (VAR_H only: the OMG entries without VAR_H are empty and never selected)
// The mixed walkers run faster below the occupancy their registers allow
// (four wavefronts per SIMD): the force walker at two (2.83 against 3.10 ms
// per launch at 160^3), the dddt walker at three (1.71 against 1.73 ms); their
// neighbour gathers thrash the caches with more wavefronts resident.  A
// dynamic LDS reservation that the kernel never touches holds a walker at
// `waves` wavefronts per SIMD, whatever its block size (64, 128 or 256
// threads): blocks + 1 of them do not fit in a 160 KB CU, blocks do.
unsigned walkerLds(const sphdem_ctx* ctx, unsigned waves) {
  const unsigned perBlock = (unsigned)ctx->walkTPB / 64u;
  const unsigned blocks = std::max(1u, 4u * waves / perBlock);
  return std::min(65536u, (163840u / (blocks + 1u) / 1024u + 1u) * 1024u);
}
constexpr unsigned kForceWaves = 2, kDddtWaves = 3;
template <int F, bool OMG>
void launchForce(sphdem_ctx* ctx) {
  constexpr bool V = (F & 1) != 0, A = (F & 2) != 0, L = (F & 4) != 0;
  constexpr int M = F >> 3;
  if constexpr (!OMG || V) {
    if (mixed(ctx))
      hipLaunchKernelGGL((k_force_m<V, M, A, L, OMG>), dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB),
                         walkerLds(ctx, kForceWaves),
                         ctx->stream, ctx->cur, ctx->ns, ctx->nbl, ctx->nbcnt, ctx->mixP, ctx->mixV,
                         ctx->mixD, ctx->mixRes, ctx->P, ctx->dg);
    else
      hipLaunchKernelGGL((k_force_l<V, M, A, L, OMG>), dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0,
                         ctx->stream, ctx->cur, ctx->ns, ctx->box, ctx->nbl, ctx->nbcnt, ctx->hinv,
                         ctx->P, ctx->dg);
  }
}
template <bool OMG, int... F>
constexpr std::array<void (*)(sphdem_ctx*), sizeof...(F)> forceTable(
    std::integer_sequence<int, F...>) {
  return {launchForce<F, OMG>...};
}
// hinvReady: the 1/h table is already current (made beside the list build);
// withOmega (VAR_H only): calcOmega summed in the same walk and omega stored --
// the step's omega pass, which only the force and dddt of omega's own particle
// read, fused into the walk that follows it
int passForce(sphdem_ctx* ctx, bool hinvReady = false, bool withOmega = false) {
  if (withOmega && !varH(ctx)) return fail(ctx, SPHDEM_E_STATE, "fused omega needs VAR_H");
  CHECK(ensureLists(ctx));
  CHECK(refreshGhosts(ctx, GS_ALL));
  if (!mixed(ctx) && !hinvReady) CHECK(computeHinv(ctx));
  static constexpr auto table = forceTable<false>(std::make_integer_sequence<int, 32>{});
  static constexpr auto tableOmega = forceTable<true>(std::make_integer_sequence<int, 32>{});
  const uint32_t f = ctx->P.flags;
  const int sel = ((f & SPHDEM_F_VAR_H_CORRECTION) ? 1 : 0) |
                  ((f & SPHDEM_F_VISC_ARTIFICIAL) ? 2 : 0) | ((f & SPHDEM_F_LIQ_DEM) ? 4 : 0) |
                  (viscVariant(f) << 3);
  Timed tm(ctx, KID_FORCE);
  if (mixed(ctx)) CHECK(packMix(ctx, MIX_P | MIX_V | MIX_D));
  (withOmega ? tableOmega : table)[sel](ctx);
  LAUNCHED();
  if (withOmega) markStale(ctx, GS_DPO);
  return 0;
}
// withOmega: as in passForce
int passDddt(sphdem_ctx* ctx, bool withOmega = false) {
  if (withOmega && !varH(ctx)) return fail(ctx, SPHDEM_E_STATE, "fused omega needs VAR_H");
  CHECK(ensureLists(ctx));
  CHECK(refreshGhosts(ctx, GS_POSH | GS_VHM));
  if (mixed(ctx)) {
    Timed tm(ctx, KID_DDDT);
    CHECK(packMix(ctx, MIX_P | MIX_V));
    if (varH(ctx) && withOmega)
      hipLaunchKernelGGL((k_dddt_m<true, true>), dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB),
                         walkerLds(ctx, kDddtWaves),
                         ctx->stream, ctx->cur, ctx->ns, ctx->nbl, ctx->nbcnt, ctx->mixP, ctx->mixV,
                         ctx->mixRes, ctx->P.wcon);
    else if (varH(ctx))
      hipLaunchKernelGGL(k_dddt_m<true>, dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0, ctx->stream,
                         ctx->cur, ctx->ns, ctx->nbl, ctx->nbcnt, ctx->mixP, ctx->mixV,
                         ctx->mixRes, ctx->P.wcon);
    else
      hipLaunchKernelGGL(k_dddt_m<false>, dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0, ctx->stream,
                         ctx->cur, ctx->ns, ctx->nbl, ctx->nbcnt, ctx->mixP, ctx->mixV,
                         ctx->mixRes, ctx->P.wcon);
  } else if (varH(ctx)) {
    Timed tm(ctx, KID_DDDT);
    if (withOmega)
      hipLaunchKernelGGL((k_dddt_l<true, true>), dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0,
                         ctx->stream, ctx->cur, ctx->ns, ctx->box, ctx->nbl, ctx->nbcnt,
                         (const double2*)nullptr, ctx->P.wcon);
    else
      hipLaunchKernelGGL(k_dddt_l<true>, dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0, ctx->stream, ctx->cur,
                         ctx->ns, ctx->box, ctx->nbl, ctx->nbcnt, (const double2*)nullptr,
                         ctx->P.wcon);
  } else {
    CHECK(computeHinv(ctx));
    Timed tm(ctx, KID_DDDT);
    hipLaunchKernelGGL(k_dddt_l<false>, dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->box, ctx->nbl, ctx->nbcnt, ctx->hinv, ctx->P.wcon);
  }
  LAUNCHED();
  if (withOmega) markStale(ctx, GS_DPO);
  return 0;
}
int passCoupling(sphdem_ctx* ctx) {
  if (ctx->nd > 0) {
    Timed tm(ctx, KID_SHEPSUM);
    hipLaunchKernelGGL(k_shepsum, dim3(waveBlocks(ctx->nd)), dim3(TPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->nd, ctx->gs, ctx->start, ctx->box, ctx->rule, ctx->reachShep,
                       ctx->P.wcon, ctx->P.coupling_radius);
    LAUNCHED();
  }
  // the DEM cells with a sphere around them (no memset: the marks carry a
  // number the device advances before each marking)
  const int cells = ctx->gc.dim[0] * ctx->gc.dim[1] * ctx->gc.dim[2];
  if (cells > ctx->demNearCap) {
    if (ctx->graphMode)
      return fail(ctx, SPHDEM_E_STATE, "a captured step cannot grow the DEM-cell marks");
    if (ctx->demNear) HIPC(hipFree(ctx->demNear));
    ctx->demNearCap = cells + cells / 4;
    HIPC(hipMalloc(&ctx->demNear, sizeof(int) * ctx->demNearCap));
    ctx->allocEpoch++;
    HIPC(hipMemsetAsync(ctx->demNear, 0, sizeof(int) * ctx->demNearCap, ctx->stream));
    HIPC(hipMemsetAsync(ctx->nearGen, 0, sizeof(int), ctx->stream));
  }
  hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, ctx->stream, ctx->nearGen);
  LAUNCHED();
  if (ctx->nd > 0) {
    hipLaunchKernelGGL(k_dem_near, dim3(blocks(ctx->nd)), dim3(TPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->nd, ctx->gc, ctx->demNear, ctx->nearGen);
    LAUNCHED();
  }
  Timed tm(ctx, KID_COUPLING);
  hipLaunchKernelGGL(k_coupling_sph, dim3(blocks(ctx->ns)), dim3(TPB), 0, ctx->stream, ctx->cur,
                     ctx->ns, ctx->nd, ctx->gc, ctx->start, ctx->box, ctx->rule, ctx->P,
                     ctx->demNear, ctx->nearGen);
  LAUNCHED();
  markStale(ctx, GS_DPO);
  return 0;
}
int passFpi(sphdem_ctx* ctx) {
  if (ctx->nd == 0) return 0;
  Timed tm(ctx, KID_FPI);
  hipLaunchKernelGGL(k_fpi, dim3(waveBlocks(ctx->nd)), dim3(TPB), 0, ctx->stream, ctx->cur, ctx->ns,
                     ctx->nd, ctx->gs, ctx->start, ctx->box, ctx->P);
  LAUNCHED();
  return 0;  // writes DEM fields only (no ghost copies them)
}
constexpr int kContactsLaneFrom = 16384;  // spheres
int passContacts(sphdem_ctx* ctx) {
  if (ctx->nd == 0) return 0;
  Timed tm(ctx, KID_CONTACTS);
  // one wavefront per sphere (four per workgroup) for the few spheres of a
  // small box, where the one-thread walk is a chain of dependent loads the
  // launch cannot hide; one thread per sphere from kContactsLaneFrom on,
  // where it runs beside the dddt walk and fits the slots that walk leaves.
  // The contact forces are summed in candidate order either way.
  if (ctx->nd >= kContactsLaneFrom)
    hipLaunchKernelGGL(k_contacts_lane, dim3((ctx->nd + 63) / 64), dim3(64), 0, ctx->stream,
                       ctx->cur, ctx->ns, ctx->nd, ctx->gc, ctx->start, ctx->box, ctx->P);
  else
    hipLaunchKernelGGL(k_contacts, dim3(waveBlocks(ctx->nd)), dim3(TPB), 0, ctx->stream,
                       ctx->cur, ctx->ns, ctx->nd, ctx->gc, ctx->start, ctx->box, ctx->P);
  LAUNCHED();
  return 0;  // writes DEM fields only
}
// one iteration of the density re-initialisation (sphIncompress.cpp:29-34)
int passDensity(sphdem_ctx* ctx) {
  hipLaunchKernelGGL(k_hfromdens, dim3(blocks(ctx->ns)), dim3(TPB), 0, ctx->stream, ctx->cur,
                     ctx->ns, ctx->P.hfac);
  LAUNCHED();
  markStale(ctx, GS_POSH);
  CHECK(ensureLists(ctx));  // the first iteration builds them with the new h
  CHECK(refreshGhosts(ctx, GS_POSH | GS_VHM));
  if (mixed(ctx)) {
    CHECK(packMix(ctx, MIX_P));
    hipLaunchKernelGGL(k_density_m, dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->nbl, ctx->nbcnt, ctx->mixP, ctx->mixRes, ctx->P.wcon);
  } else {
    CHECK(packPosm(ctx));
    hipLaunchKernelGGL(k_density_l, dim3(walkBlocks(ctx, ctx->ns)), dim3(ctx->walkTPB), 0, ctx->stream, ctx->cur,
                       ctx->ns, ctx->box, ctx->nbl, ctx->nbcnt, ctx->posm, ctx->P.wcon);
  }
  LAUNCHED();
  markStale(ctx, GS_DPO);
  return 0;
}

// ---- half-steps
// whether start() of step `step` re-initialises the density (sphIncompress.cpp:17-35)
bool reinitDue(const sphdem_ctx* c, long long step) {
  return (step % c->P.reinit_every == 0) || (step == c->P.reinit_at);
}
// bodyMask/accel: the beforeMiddle body-force hook, applied by k_start right
// after initSumsMiddle zeroes f (nothing in between touches f)
int stepStart(sphdem_ctx* ctx, uint32_t bodyMask = 0, const double* accel = nullptr) {
  const int na = ctx->ns + ctx->nd;
  const double a0 = accel ? accel[0] : 0.0, a1 = accel ? accel[1] : 0.0,
               a2 = accel ? accel[2] : 0.0;
  {
    Timed tm(ctx, KID_START);
    hipLaunchKernelGGL(k_start, dim3(blocks(std::max(na, 1))), dim3(TPB), 0, ctx->stream,
                       ctx->cur, na, ctx->P, ctx->dg, accel ? bodyMask : 0u, a0, a1, a2);
    LAUNCHED();
    markStale(ctx, GS_ALL);
  }
  if (reinitDue(ctx, ctx->sphStep)) {
    CHECK(doReset(ctx));
    for (int i = 0; i < 10; i++) CHECK(passDensity(ctx));
  }
  return 0;
}
// fold: start() (WrapFold mode 1) when the step's start() has not run
int stepMiddle(sphdem_ctx* ctx, const WrapFold& fold = WrapFold{}) {
  CHECK(doReset(ctx, false, true, fold));
  bool hinvReady = false;
  {
    // On the side stream, beside the list build: the coupling pass (porosity,
    // Shepard sums, DEM->SPH drag), then the EOS (reads dens and porosity,
    // stores pdr2 and press) and the FP64 force pass's 1/h table.  None of
    // them writes what the list build reads.  (The ghosts are fresh after the
    // reset, so nothing on the side stream refreshes them.)
    CHECK(refreshGhosts(ctx, GS_POSH));
    Side side(ctx);
    if (liqDem(ctx)) CHECK(passCoupling(ctx));
    CHECK(passEos(ctx));
    if (!mixed(ctx)) {
      CHECK(computeHinv(ctx));
      hinvReady = true;
    }
    side.back();
    CHECK(ensureLists(ctx));
    CHECK(side.join());
  }
  // calcOmega (VAR_H) is summed in the force walk: the force of a reads omega_a only
  CHECK(passForce(ctx, hinvReady, varH(ctx)));
  const int na = ctx->ns + ctx->nd;
  hipLaunchKernelGGL(k_accel, dim3(reduceBlocks(std::max(na, 1))), dim3(TPB), 0, ctx->stream,
                     ctx->cur, na, ctx->dg);
  LAUNCHED();
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1), 0, ctx->stream, ctx->P, ctx->dg);
  LAUNCHED();
  Timed tm(ctx, KID_KICK);
  hipLaunchKernelGGL(k_mid_kick, dim3(blocks(std::max(na, 1))), dim3(TPB), 0, ctx->stream,
                     ctx->cur, na, ctx->P, ctx->dg);
  LAUNCHED();
  markStale(ctx, GS_ALL);
  return 0;
}
// foldDrift: the drift of r (and the time's half step) done by the reset's k_wrap
int stepEnd(sphdem_ctx* ctx, bool foldDrift = false) {
  WrapFold fold{};
  if (foldDrift) {
    fold.mode = 2;
    fold.nlive = ctx->ns;
  } else {
    hipLaunchKernelGGL(k_end_drift, dim3(blocks(std::max(ctx->ns, 1))), dim3(TPB), 0, ctx->stream,
                       ctx->cur, ctx->ns, ctx->dg);
    LAUNCHED();
    markStale(ctx, GS_ALL);
  }
  CHECK(doReset(ctx, false, false, fold));
  // calcOmega (VAR_H) is summed in the dddt walk: the dddt of a reads omega_a only
  const bool fuse = varH(ctx);
  if (liqDem(ctx)) {
    // calcFPI, DEM contacts and the DEM kick on the side stream, beside dddt:
    // they read SPH positions, h, dens, vhat, porosity and forces and write DEM
    // fields only; dddt writes the SPH dddt (and omega, which they do not read)
    // only.  driftRest (k_end_rest) changes SPH dens and h, so it waits for both.
    Side side(ctx);
    CHECK(passFpi(ctx));
    if (!(ctx->P.flags & (SPHDEM_F_FIXED_DEM | SPHDEM_F_NO_DEM_CONTACTS))) CHECK(passContacts(ctx));
    // calcDEMWallContacts (without LIQ_DEM_CUSTOM_WALL_CONTACTS) scans the
    // DEM's own neighbour list, which holds DEM particles only (DEM are binned
    // in dem_cells alone, dataLL.cpp:979-990), for boundary particles: it never
    // finds one and adds nothing, so there is no kernel for it
    if ((ctx->P.flags & SPHDEM_F_DEM_BLOCK) && ctx->nd > 0) {
      hipLaunchKernelGGL(k_dem_block, dim3(1), dim3(kDemBlockThreads), 0, ctx->stream, ctx->cur,
                         ctx->ns, ctx->nd);
      LAUNCHED();
    }
    if (ctx->nd > 0) {
      hipLaunchKernelGGL(k_dem_kick, dim3(blocks(ctx->nd)), dim3(TPB), 0, ctx->stream, ctx->cur,
                         ctx->ns, ctx->nd, ctx->P, ctx->dg);
      LAUNCHED();
    }
    side.back();
    CHECK(passDddt(ctx, fuse));
    CHECK(side.join());
  } else {
    CHECK(passDddt(ctx, fuse));
  }
  Timed tm(ctx, KID_REST);
  hipLaunchKernelGGL(k_end_rest, dim3(blocks(std::max(ctx->ns, 1))), dim3(TPB), 0, ctx->stream,
                     ctx->cur, ctx->ns, ctx->P, ctx->dg);
  LAUNCHED();
  markStale(ctx, GS_ALL);
  return 0;
}

int hookBody(sphdem_ctx* ctx, uint32_t mask, const double* a) {
  const int na = ctx->ns + ctx->nd;
  hipLaunchKernelGGL(k_body_force, dim3(blocks(std::max(na, 1))), dim3(TPB), 0, ctx->stream,
                     ctx->cur, na, mask, a[0], a[1], a[2]);
  LAUNCHED();
  markStale(ctx, GS_ALL);
  return 0;
}
int hookFloor(sphdem_ctx* ctx) {
  if (ctx->nd == 0) return 0;
  hipLaunchKernelGGL(k_dem_floor, dim3(blocks(ctx->nd)), dim3(TPB), 0, ctx->stream, ctx->cur,
                     ctx->ns, ctx->nd, ctx->P);
  LAUNCHED();
  markStale(ctx, GS_ALL);
  return 0;
}
int hookFreeze(sphdem_ctx* ctx, uint32_t mask, double t) {
  const int na = ctx->ns + ctx->nd;
  hipLaunchKernelGGL(k_freeze, dim3(blocks(std::max(na, 1))), dim3(TPB), 0, ctx->stream, ctx->cur,
                     na, mask, t, ctx->dg);
  LAUNCHED();
  markStale(ctx, GS_ALL);
  return 0;
}
int hookSumDrag(sphdem_ctx* ctx) {
  const int na = std::max(ctx->ns + ctx->nd, 1);
  const int nb = blocks(na);
  if (nb * 6 > ctx->partialCap) {
    if (ctx->partial) HIPC(hipFree(ctx->partial));
    ctx->partialCap = nb * 6;
    HIPC(hipMalloc(&ctx->partial, sizeof(double) * ctx->partialCap));
    ctx->allocEpoch++;
  }
  hipLaunchKernelGGL(k_sum_drag, dim3(nb), dim3(TPB), 0, ctx->stream, ctx->cur, ctx->ns + ctx->nd,
                     ctx->partial);
  LAUNCHED();
  hipLaunchKernelGGL(k_sum_drag_final, dim3(1), dim3(256), 0, ctx->stream, ctx->partial, nb,
                     ctx->dg);
  LAUNCHED();
  return 0;
}

void freeAll(sphdem_ctx* ctx) {
  void* ptrs[] = {ctx->storeMem[0], ctx->storeMem[1], ctx->key,     ctx->slot,   ctx->tmp,
                  ctx->perm,        ctx->counts,      ctx->start,   ctx->tiles,
                  ctx->partial,     ctx->fieldbuf, ctx->errflag, ctx->dg,
                  ctx->segCnt,      ctx->segOff,      ctx->segs,    ctx->nbl,    ctx->nbcnt,
                  ctx->hinv,        ctx->gcnt,        ctx->goff,    ctx->gsrc,   ctx->gord,
                  ctx->gcell,       ctx->gorig,       ctx->gcode,   ctx->inv,
                  ctx->mixP,        ctx->mixV,        ctx->mixD,    ctx->okey,
                  ctx->posm,        ctx->pairCnt,     ctx->pairOff, ctx->lrec,
                  ctx->lflag,       ctx->demNear,     ctx->nearGen, ctx->gate,
                  ctx->keepMem,     ctx->dgKeep};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : ctx->evPool) (void)hipEventDestroy(e);
  ctx->evPool.clear();
  if (ctx->hostG) (void)hipHostFree(ctx->hostG);
  ctx->hostG = nullptr;
  if (ctx->gateHost) (void)hipHostFree(ctx->gateHost);
  ctx->gateHost = nullptr;
  if (ctx->graphExec) (void)hipGraphExecDestroy(ctx->graphExec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->graphExec = nullptr;
  ctx->graph = nullptr;
  if (ctx->evFork) (void)hipEventDestroy(ctx->evFork);
  if (ctx->evJoin) (void)hipEventDestroy(ctx->evJoin);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

}  // namespace

// ============================================================ C ABI
extern "C" {

// at most one viscosity variant (calcViscForce, sph.h:579-631)
static bool oneViscosity(uint32_t f) {
  const uint32_t v = f & (SPHDEM_F_VISC_MORRIS | SPHDEM_F_VISC_CLEARY | SPHDEM_F_VISC_MONAGHAN);
  return (v & (v - 1)) == 0;
}

int sphdem_create(const sphdem_params* params, int device, int64_t capacity, sphdem_ctx** out) {
  if (!params || !out || capacity <= 0 || capacity > (int64_t)0x3fffffff) return SPHDEM_E_ARG;
  if (!oneViscosity(params->flags)) return SPHDEM_E_ARG;
  *out = nullptr;
  sphdem_ctx* ctx = new sphdem_ctx();
  ctx->P = *params;
  ctx->device = device;
  ctx->cap = capacity;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    delete ctx;
    return SPHDEM_E_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    delete ctx;
    return SPHDEM_E_DEVICE;
  }
  int rc = 0;
  auto setup = [&]() -> int {
    HIPC(hipSetDevice(device));
    HIPC(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    CHECK(allocEntries(ctx, capacity / 4 + 4096));
    HIPC(hipMalloc(&ctx->gcnt, sizeof(int) * (capacity + 1)));
    HIPC(hipMalloc(&ctx->goff, sizeof(int) * (capacity + 1)));
    HIPC(hipMalloc(&ctx->inv, sizeof(int) * capacity));
    HIPC(hipMalloc(&ctx->nbl, sizeof(int) * (size_t)kMaxNb * capacity));
    HIPC(hipMalloc(&ctx->nbcnt, sizeof(int) * capacity));
    HIPC(hipMalloc(&ctx->fieldbuf, sizeof(double) * 3 * capacity));
    HIPC(hipMalloc(&ctx->errflag, sizeof(int)));
    HIPC(hipMalloc(&ctx->lflag, sizeof(int) * 2));
    HIPC(hipMemset(ctx->lflag, 0, sizeof(int) * 2));
    HIPC(hipMalloc(&ctx->nearGen, sizeof(int)));
    HIPC(hipMemset(ctx->nearGen, 0, sizeof(int)));
    HIPC(hipMalloc(&ctx->gate, sizeof(int) * 4));
    HIPC(hipMemset(ctx->gate, 0, sizeof(int) * 4));
    HIPC(hipHostMalloc((void**)&ctx->gateHost, sizeof(int) * 4, hipHostMallocDefault));
    HIPC(hipMalloc(&ctx->dg, sizeof(DevGlobals)));
    HIPC(hipMemset(ctx->dg, 0, sizeof(DevGlobals)));
    HIPC(hipHostMalloc((void**)&ctx->hostG, sizeof(DevGlobals), hipHostMallocDefault));
    return 0;
  };
  rc = setup();
  if (rc) {
    freeAll(ctx);
    delete ctx;
    return rc;
  }
  setupStatic(ctx);
  *out = ctx;
  return 0;
}

int sphdem_destroy(sphdem_ctx* ctx) {
  if (!ctx) return SPHDEM_E_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  freeAll(ctx);
  delete ctx;
  return 0;
}

const char* sphdem_last_error(const sphdem_ctx* ctx) { return ctx ? ctx->err : "null context"; }

int sphdem_set_params(sphdem_ctx* ctx, const sphdem_params* params) {
  if (!ctx || !params) return SPHDEM_E_ARG;
  if (!oneViscosity(params->flags))
    return fail(ctx, SPHDEM_E_ARG, "more than one of VISC_MORRIS, VISC_CLEARY, VISC_MONAGHAN");
  for (int d = 0; d < 3; d++)
    if (params->periodic[d] != ctx->P.periodic[d] || params->proc_domain[2 * d] != ctx->P.proc_domain[2 * d] ||
        params->proc_domain[2 * d + 1] != ctx->P.proc_domain[2 * d + 1])
      return fail(ctx, SPHDEM_E_ARG, "the domain cannot change on a live context");
  ctx->P = *params;
  setupStatic(ctx);
  return 0;
}

int sphdem_upload(sphdem_ctx* ctx, int64_t n, const double* r, const double* v, const double* vhat,
                  const double* mass, const double* h, const double* dens, const int32_t* iam) {
  if (!ctx || n < 0 || !r || !mass || !h || !dens || !iam) return SPHDEM_E_ARG;
  if (n > ctx->cap) return fail(ctx, SPHDEM_E_CAPACITY, "%lld particles exceed capacity %lld",
                                (long long)n, (long long)ctx->cap);
  HIPC(hipSetDevice(ctx->device));
  std::vector<double4> a[kNumVec];
  for (auto& x : a) x.assign((size_t)n, make_double4(0, 0, 0, 0));
  std::vector<int> t((size_t)n), o((size_t)n);
  int ns = 0, nd = 0;
  for (int64_t i = 0; i < n; i++) {
    const int ty = iam[i];
    if (ty == SPHDEM_SPH || ty == SPHDEM_SPH_BOUNDARY)
      ns++;
    else if (ty == SPHDEM_DEM)
      nd++;
    else
      return fail(ctx, SPHDEM_E_ARG,
                  "particle %lld has type %d; supported types are sph, sphBoundary and dem",
                  (long long)i, ty);
    a[0][i] = make_double4(r[3 * i], r[3 * i + 1], r[3 * i + 2], h[i]);
    a[1][i] = vhat ? make_double4(vhat[3 * i], vhat[3 * i + 1], vhat[3 * i + 2], mass[i])
                   : make_double4(0, 0, 0, mass[i]);
    a[2][i] = make_double4(dens[i], 0.0, 0.0, 1.0);  // omega 0, porosity 1 (particle.h:77,80)
    if (v) a[3][i] = make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.0);
    t[i] = ty;
    o[i] = (int)i;
  }
  for (int k = 0; k < kNumVec; k++)
    HIPC(hipMemcpy(storeArr(ctx->cur, k), a[k].data(), sizeof(double4) * n, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(ctx->cur.iam, t.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(ctx->cur.orig, o.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  ctx->n = (int)n;
  ctx->ns0 = ns;
  ctx->nd0 = nd;
  ctx->ns = ns;
  ctx->nd = nd;
  DevGlobals g;
  memset(&g, 0, sizeof(g));
  g.maxdt = ctx->P.maxdt;
  HIPC(hipMemcpy(ctx->dg, &g, sizeof(g), hipMemcpyHostToDevice));
  ctx->sphStep = 0;
  // CdataLL::CdataLL (dataLL.cpp:8-40): hmax = sph_search_radius = H, then reset()
  ctx->ssr = ctx->P.h;
  ctx->gridValid = false;
  ctx->loaded = true;
  CHECK(doReset(ctx));
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int sphdem_num_particles(const sphdem_ctx* ctx, int64_t* n) {
  if (!ctx || !n) return SPHDEM_E_ARG;
  *n = ctx->n;
  return 0;
}

int sphdem_output_vars(sphdem_ctx* ctx, sphdem_output* out) {
  CHECK(checkLoaded(ctx));
  if (!out) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  const int ns = ctx->ns;
  CHECK(ensureLists(ctx));
  CHECK(checkFlags(ctx));
  CHECK(refreshGhosts(ctx, GS_POSH | GS_VHM));
  double* densCalc = ctx->fieldbuf;
  if (ns > 0) {
    CHECK(packPosm(ctx));
    hipLaunchKernelGGL(k_density_l, dim3(walkBlocks(ctx, ns)), dim3(ctx->walkTPB), 0, ctx->stream, ctx->cur, ns,
                       ctx->box, ctx->nbl, ctx->nbcnt, ctx->posm, ctx->P.wcon, densCalc);
    LAUNCHED();
  }
  const int nb = std::max(blocks(ns), 1);
  if (nb * OV_COUNT + OV_COUNT > ctx->partialCap) {
    if (ctx->partial) HIPC(hipFree(ctx->partial));
    ctx->partialCap = nb * OV_COUNT + OV_COUNT;
    HIPC(hipMalloc(&ctx->partial, sizeof(double) * ctx->partialCap));
    ctx->allocEpoch++;
  }
  double* fin = ctx->partial + (size_t)nb * OV_COUNT;
  double h0[OV_COUNT], h1[OV_COUNT];
  auto run = [&](int pass, double ave, double* host) -> int {
    hipLaunchKernelGGL(k_output_partials, dim3(nb), dim3(256), 0, ctx->stream, ctx->cur, ns,
                       ctx->P, densCalc, ave, pass, ctx->partial);
    LAUNCHED();
    hipLaunchKernelGGL(k_output_final, dim3(1), dim3(64), 0, ctx->stream, ctx->partial, nb, fin);
    LAUNCHED();
    HIPC(hipMemcpyAsync(host, fin, sizeof(double) * OV_COUNT, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return 0;
  };
  CHECK(run(0, 0.0, h0));
  const double n = (double)(ctx->ns + ctx->nd);  // the container's particles
  const double aveDens = n > 0 ? h0[OV_AVEDENS] / n : 0.0;
  CHECK(run(1, aveDens, h1));
  sphdem_output o = {};
  o.linMom[0] = h0[OV_LINX];
  o.linMom[1] = h0[OV_LINY];
  o.linMom[2] = h0[OV_LINZ];
  o.angMom = h0[OV_ANG];
  o.eKE = h0[OV_EKE];
  o.eViscF = h0[OV_EVISCF];
  o.edViscFdt = h0[OV_EDVISCFDT];
  o.eElast = h0[OV_EELAST];
  o.edElastdt = h0[OV_EDELASTDT];
  o.eElastExact = h0[OV_EELASTEXACT];
  o.maxV = h0[OV_MAXV];
  o.maxF = h0[OV_MAXF];
  o.aveDens = aveDens;
  o.aveDensFromMass = n > 0 ? sqrt(h0[OV_DENSFROMMASS] / n) : 0.0;
  o.varDens = n > 0 ? h1[OV_VARDENS] / n : 0.0;
  o.n = (int64_t)n;
  o.nSph = (int64_t)h0[OV_NSPH];
  o.eTotal = o.eKE + o.eElast + o.eViscF + o.eViscB + o.eBForce + o.eFF;
  *out = o;
  return 0;
}

// ---- full-state checkpoint --------------------------------------------------------
namespace {
struct CheckpointHeader {
  char magic[8];  // "SPHDEMCK"
  int64_t version, n, ns0, nd0, ns, nd, sphStep;
  double ssr;
  int64_t paramsBytes, globalsBytes;
};
constexpr int64_t kCheckpointVersion = 1;
int64_t checkpointBytes(int64_t n) {
  return (int64_t)sizeof(CheckpointHeader) + (int64_t)sizeof(sphdem_params) +
         (int64_t)sizeof(DevGlobals) + n * (int64_t)(kNumVec * sizeof(double4) + sizeof(int));
}
}  // namespace

int sphdem_checkpoint_size(const sphdem_ctx* ctx, int64_t* bytes) {
  if (!ctx || !bytes) return SPHDEM_E_ARG;
  *bytes = checkpointBytes(ctx->n);
  return 0;
}

int sphdem_checkpoint_save(sphdem_ctx* ctx, void* host, int64_t capacity) {
  CHECK(checkLoaded(ctx));
  if (!host) return SPHDEM_E_ARG;
  const int64_t need = checkpointBytes(ctx->n);
  if (capacity < need)
    return fail(ctx, SPHDEM_E_CAPACITY, "checkpoint needs %lld bytes", (long long)need);
  HIPC(hipSetDevice(ctx->device));
  const int n = ctx->n;
  char* p = (char*)host;
  CheckpointHeader h;
  memcpy(h.magic, "SPHDEMCK", 8);
  h.version = kCheckpointVersion;
  h.n = n;
  h.ns0 = ctx->ns0;
  h.nd0 = ctx->nd0;
  h.ns = ctx->ns;
  h.nd = ctx->nd;
  h.sphStep = ctx->sphStep;
  h.ssr = ctx->ssr;
  h.paramsBytes = sizeof(sphdem_params);
  h.globalsBytes = sizeof(DevGlobals);
  memcpy(p, &h, sizeof(h));
  p += sizeof(h);
  memcpy(p, &ctx->P, sizeof(sphdem_params));
  p += sizeof(sphdem_params);
  HIPC(hipMemcpyAsync(p, ctx->dg, sizeof(DevGlobals), hipMemcpyDeviceToHost, ctx->stream));
  p += sizeof(DevGlobals);
  // each stream in original-index order, staged through the other store buffer
  double4* stage = ctx->alt.posh;
  for (int k = 0; k < kNumVec; k++) {
    hipLaunchKernelGGL(k_gather_orig, dim3(blocks(n)), dim3(TPB), 0, ctx->stream,
                       storeArr(ctx->cur, k), ctx->cur.orig, n, stage);
    LAUNCHED();
    HIPC(hipMemcpyAsync(p, stage, sizeof(double4) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    p += sizeof(double4) * n;
  }
  hipLaunchKernelGGL(k_gather_orig_int, dim3(blocks(n)), dim3(TPB), 0, ctx->stream, ctx->cur.iam,
                     ctx->cur.orig, n, ctx->alt.iam);
  LAUNCHED();
  HIPC(hipMemcpyAsync(p, ctx->alt.iam, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int sphdem_checkpoint_restore(sphdem_ctx* ctx, const void* host, int64_t bytes) {
  if (!ctx || !host || bytes < (int64_t)sizeof(CheckpointHeader)) return SPHDEM_E_ARG;
  CheckpointHeader h;
  memcpy(&h, host, sizeof(h));
  if (memcmp(h.magic, "SPHDEMCK", 8) != 0 || h.version != kCheckpointVersion ||
      h.paramsBytes != (int64_t)sizeof(sphdem_params) ||
      h.globalsBytes != (int64_t)sizeof(DevGlobals) || h.n < 0 || bytes != checkpointBytes(h.n))
    return fail(ctx, SPHDEM_E_ARG, "not a checkpoint of this library version");
  if (h.n > ctx->cap)
    return fail(ctx, SPHDEM_E_CAPACITY, "%lld particles exceed capacity %lld", (long long)h.n,
                (long long)ctx->cap);
  const char* p = (const char*)host + sizeof(h);
  if (memcmp(p, &ctx->P, sizeof(sphdem_params)) != 0)
    return fail(ctx, SPHDEM_E_ARG, "the checkpoint was taken with other parameters");
  p += sizeof(sphdem_params);
  HIPC(hipSetDevice(ctx->device));
  // the hmax reduction target starts clear, whatever the saved value
  DevGlobals dgh;
  memcpy(&dgh, p, sizeof(DevGlobals));
  dgh.hmax_bits = 0;
  HIPC(hipMemcpyAsync(ctx->dg, &dgh, sizeof(DevGlobals), hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));  // dgh is a local
  p += sizeof(DevGlobals);
  const int n = (int)h.n;
  for (int k = 0; k < kNumVec; k++) {
    HIPC(hipMemcpyAsync(storeArr(ctx->cur, k), p, sizeof(double4) * n, hipMemcpyHostToDevice,
                        ctx->stream));
    p += sizeof(double4) * n;
  }
  HIPC(hipMemcpyAsync(ctx->cur.iam, p, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_identity, dim3(blocks(std::max(n, 1))), dim3(TPB), 0, ctx->stream,
                     ctx->cur.orig, n);
  LAUNCHED();
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->n = n;
  ctx->ns0 = (int)h.ns0;
  ctx->nd0 = (int)h.nd0;
  ctx->ns = (int)h.ns;
  ctx->nd = (int)h.nd;
  ctx->sphStep = h.sphStep;
  ctx->ssr = h.ssr;
  ctx->ng = 0;
  ctx->gridValid = false;
  ctx->listsValid = false;
  ctx->loaded = true;
  // the store is in original-index order: sort it with the saved search
  // radius (the grid of the saved state)
  return doReset(ctx, true);
}

int sphdem_set_tile_shape(sphdem_ctx* ctx, int seg_particles, int walker_block) {
  if (!ctx) return SPHDEM_E_ARG;
  if ((seg_particles != 0 && (seg_particles < 8 || seg_particles > 64)) ||
      (walker_block != 64 && walker_block != 128 && walker_block != 256))
    return fail(ctx, SPHDEM_E_ARG,
                "tile shape %d particles per segment, %d threads per walker block: segments "
                "hold 8..64 particles (0: per-target search), walker blocks 64, 128 or 256 "
                "threads",
                seg_particles, walker_block);
  // the segments of the last reset stay a valid partition until the next
  // reset rebuilds them with the new size; after a reset without segments
  // (the per-target search) the next segmented build makes them
  ctx->tileSeg = seg_particles;
  ctx->walkTPB = walker_block;
  return 0;
}

int sphdem_get_tile_shape(const sphdem_ctx* ctx, int* seg_particles, int* walker_block) {
  if (!ctx || !seg_particles || !walker_block) return SPHDEM_E_ARG;
  *seg_particles = ctx->tileSeg;
  *walker_block = ctx->walkTPB;
  return 0;
}

int sphdem_tune_tile_shape(sphdem_ctx* ctx, int steps, int hookset, int* seg_particles,
                           int* walker_block, double* ms_per_step) {
  CHECK(checkLoaded(ctx));
  if (steps < 1 || steps > 1000 || !seg_particles || !walker_block) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  int64_t bytes = 0;
  CHECK(sphdem_checkpoint_size(ctx, &bytes));
  std::vector<char> blob((size_t)bytes);
  CHECK(sphdem_checkpoint_save(ctx, blob.data(), bytes));
  static const int segs[] = {0, 64, 48, 32};
  static const int tpbs[] = {256, 128, 64};
  const int seg0 = ctx->tileSeg, tpb0 = ctx->walkTPB;
  double best = INFINITY;
  int bestSeg = seg0, bestTpb = tpb0;
  int rc = 0;
  for (int sg : segs) {
    for (int tb : tpbs) {
      ctx->tileSeg = sg;
      ctx->walkTPB = tb;
      // the restore's reset rebuilds the segments with this size
      if ((rc = sphdem_checkpoint_restore(ctx, blob.data(), bytes)) != 0) break;
      if ((rc = sphdem_run_steps(ctx, 1, hookset)) != 0) break;  // first launches untimed
      if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        rc = SPHDEM_E_HIP;
        break;
      }
      const auto t0 = std::chrono::steady_clock::now();
      if ((rc = sphdem_run_steps(ctx, steps, hookset)) != 0) break;
      if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        rc = SPHDEM_E_HIP;
        break;
      }
      const double el =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (el < best) {
        best = el;
        bestSeg = sg;
        bestTpb = tb;
      }
    }
    if (rc) break;
  }
  ctx->tileSeg = rc ? seg0 : bestSeg;
  ctx->walkTPB = rc ? tpb0 : bestTpb;
  const int rrc = sphdem_checkpoint_restore(ctx, blob.data(), bytes);
  if (rc) return rc;
  if (rrc) return rrc;
  *seg_particles = bestSeg;
  *walker_block = bestTpb;
  if (ms_per_step) *ms_per_step = 1e3 * best / steps;
  return 0;
}

int sphdem_num_deleted(const sphdem_ctx* ctx, int64_t* n) {
  if (!ctx || !n) return SPHDEM_E_ARG;
  *n = (int64_t)(ctx->ns0 - ctx->ns) + (ctx->nd0 - ctx->nd);
  return 0;
}

int sphdem_get_field(sphdem_ctx* ctx, int field, void* host) {
  CHECK(checkLoaded(ctx));
  if (!host) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  const int n = ctx->n;
  if (field == SPHDEM_FIELD_IAM || field == SPHDEM_FIELD_TAG) {
    int* buf = (int*)ctx->fieldbuf;
    hipLaunchKernelGGL(k_get_int, dim3(blocks(n)), dim3(TPB), 0, ctx->stream,
                       field == SPHDEM_FIELD_IAM ? ctx->cur.iam : ctx->cur.orig,
                       field == SPHDEM_FIELD_IAM ? 0 : 1, ctx->cur.orig, n, buf);
    LAUNCHED();
    HIPC(hipMemcpyAsync(host, buf, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return 0;
  }
  FieldLoc L;
  if (!fieldLoc(field, &L)) return fail(ctx, SPHDEM_E_ARG, "unknown field %d", field);
  hipLaunchKernelGGL(k_get_field, dim3(blocks(n)), dim3(TPB), 0, ctx->stream,
                     (const double*)storeArr(ctx->cur, L.arr), L.off, L.width, ctx->cur.orig, n,
                     ctx->fieldbuf);
  LAUNCHED();
  HIPC(hipMemcpyAsync(host, ctx->fieldbuf, sizeof(double) * L.width * n, hipMemcpyDeviceToHost,
                      ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

// src: n values (n*3 for vectors) in original order, in host memory or
// (device) in device memory of the context's device
static int setField(sphdem_ctx* ctx, int field, const void* src, bool device) {
  CHECK(checkLoaded(ctx));
  if (!src) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  const int n = ctx->n;
  if (field == SPHDEM_FIELD_IAM) {
    if (device) return fail(ctx, SPHDEM_E_ARG, "particle types are fixed at upload");
    const int32_t* t = (const int32_t*)src;
    for (int i = 0; i < n; i++)
      if (t[i] != SPHDEM_SPH && t[i] != SPHDEM_SPH_BOUNDARY && t[i] != SPHDEM_DEM &&
          t[i] != SPHDEM_DELETED)
        return fail(ctx, SPHDEM_E_ARG, "unsupported particle type %d", t[i]);
    return fail(ctx, SPHDEM_E_ARG, "particle types are fixed at upload");
  }
  FieldLoc L;
  if (!fieldLoc(field, &L)) return fail(ctx, SPHDEM_E_ARG, "unknown or read-only field %d", field);
  const double* from = (const double*)src;
  if (!device) {
    HIPC(hipMemcpyAsync(ctx->fieldbuf, src, sizeof(double) * L.width * n, hipMemcpyHostToDevice,
                        ctx->stream));
    from = (const double*)ctx->fieldbuf;
  }
  hipLaunchKernelGGL(k_set_field, dim3(blocks(n)), dim3(TPB), 0, ctx->stream,
                     (double*)storeArr(ctx->cur, L.arr), L.off, L.width, ctx->cur.orig, n, from);
  LAUNCHED();
  markStale(ctx, GS_ALL);
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int sphdem_set_field(sphdem_ctx* ctx, int field, const void* host) {
  return setField(ctx, field, host, false);
}

int sphdem_set_field_device(sphdem_ctx* ctx, int field, const void* dev) {
  return setField(ctx, field, dev, true);
}

int sphdem_set_globals(sphdem_ctx* ctx, const sphdem_globals* g) {
  if (!ctx || !g) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  DevGlobals d;
  HIPC(hipMemcpyAsync(&d, ctx->dg, sizeof(d), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  d.time = g->time;
  d.dt = g->dt;
  d.newDt = g->newDt;
  d.maxdt = g->maxdt;
  d.sphStep = g->sphStep;
  memcpy(&d.newDt_bits, &g->newDt, sizeof(double));
  for (int i = 0; i < 9; i++) d.custom[i] = g->custom[i];
  HIPC(hipMemcpyAsync(ctx->dg, &d, sizeof(d), hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->sphStep = g->sphStep;
  return 0;
}

int sphdem_get_globals(sphdem_ctx* ctx, sphdem_globals* g) {
  if (!ctx || !g) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  DevGlobals d;
  HIPC(hipMemcpyAsync(&d, ctx->dg, sizeof(d), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  g->time = d.time;
  g->dt = d.dt;
  g->newDt = d.newDt;
  g->maxdt = d.maxdt;
  g->sphStep = ctx->sphStep;
  for (int i = 0; i < 9; i++) g->custom[i] = d.custom[i];
  return 0;
}

int sphdem_sync(sphdem_ctx* ctx) {
  if (!ctx) return SPHDEM_E_ARG;
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int sphdem_step_start(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  CHECK(stepStart(ctx));
  return checkFlags(ctx);
}
int sphdem_step_middle(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  CHECK(stepMiddle(ctx));
  return checkFlags(ctx);
}
int sphdem_step_end(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  CHECK(stepEnd(ctx));
  return checkFlags(ctx);
}

int sphdem_hook_body_force(sphdem_ctx* ctx, uint32_t type_mask, const double accel[3]) {
  CHECK(checkLoaded(ctx));
  if (!accel) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  return hookBody(ctx, type_mask, accel);
}
int sphdem_hook_dem_floor(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  return hookFloor(ctx);
}
int sphdem_hook_freeze(sphdem_ctx* ctx, uint32_t type_mask, double t_freeze) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  return hookFreeze(ctx, type_mask, t_freeze);
}
int sphdem_hook_sum_drag(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  return hookSumDrag(ctx);
}

namespace {
// nsteps of run.cpp:102-118; the error bits of the last step's lists are left
// to the caller's checkFlags (each reset reads those of the step before)
int runOneStep(sphdem_ctx* ctx, int hookset) {
  const double g[3] = {0.0, 0.0, -9.81};
  const uint32_t sphMask = 1u << SPHDEM_SPH, demMask = 1u << SPHDEM_DEM;
  {
    ctx->sphStep++;
    // start(), then the beforeMiddle gravity on sph (fused into k_start); on
    // a step without a density re-initialisation nothing reads start()'s
    // results before middle()'s reset, which does start() in its own passes
    const bool body = hookset != SPHDEM_HOOKS_NONE;
    if (reinitDue(ctx, ctx->sphStep)) {
      if (body)
        CHECK(stepStart(ctx, sphMask, g));
      else
        CHECK(stepStart(ctx));
      CHECK(stepMiddle(ctx));
    } else {
      WrapFold fold{};
      fold.mode = 1;
      fold.nlive = ctx->ns + ctx->nd;
      fold.flags = ctx->P.flags;
      fold.hfac = ctx->P.hfac;
      if (body) {
        fold.bodyMask = sphMask;
        fold.ax = g[0];
        fold.ay = g[1];
        fold.az = g[2];
      }
      CHECK(stepMiddle(ctx, fold));
    }
    if (hookset == SPHDEM_HOOKS_SEDIMENT) {  // beforeEnd, sediment3D/customSim.cpp:97-107
      CHECK(hookSumDrag(ctx));
      // gravity, the floor contact and the freeze on dem, fused per particle
      if (ctx->nd > 0) {
        hipLaunchKernelGGL(k_hooks_dem_end, dim3(blocks(ctx->nd)), dim3(TPB), 0, ctx->stream,
                           ctx->cur, ctx->ns, ctx->nd, g[0], g[1], g[2], ctx->P,
                           ctx->P.time_drop, ctx->dg);
        LAUNCHED();
        markStale(ctx, GS_ALL);
      }
    }
    CHECK(stepEnd(ctx, true));
    if (hookset == SPHDEM_HOOKS_SEDIMENT) CHECK(hookFreeze(ctx, demMask, ctx->P.time_drop));
  }
  // every buffer a step uses is allocated now (device resets allocate nothing)
  if (!ctx->graphMode) ctx->readyEpoch = ctx->allocEpoch;
  return 0;
}
int runSteps(sphdem_ctx* ctx, int nsteps, int hookset) {
  CHECK(checkLoaded(ctx));
  if (nsteps < 0 || hookset < SPHDEM_HOOKS_NONE || hookset > SPHDEM_HOOKS_SEDIMENT)
    return fail(ctx, SPHDEM_E_ARG, "bad nsteps/hookset");
  HIPC(hipSetDevice(ctx->device));
  for (int s = 0; s < nsteps; s++) CHECK(runOneStep(ctx, hookset));
  return 0;
}

// ---- steps replayed from a captured graph ----------------------------------------
// One step (start, hooks, middle, end) is captured with the resets in their
// device form: the host reads nothing back, the ghost count stays on the
// device (bounded by graphG entries) and k_reset_gate makes the checks the
// host would make.  Each replay first keeps a copy of the state it starts
// from (k_snapshot).  A replay whose reset meets what only the host handles
// (a grid change, deletions, more ghosts, an error bit) sets the gate; the
// gate stays set, so later replays keep no copy, and after the replays the
// host restores the copy and runs that step as launches.
HostFields saveHost(const sphdem_ctx* c) {
  HostFields h;
  h.sphStep = c->sphStep;
  h.ng = c->ng;
  h.ns = c->ns;
  h.nd = c->nd;
  h.K = c->K;
  h.lgen = c->lgen;
  h.countsClean = c->countsClean;
  h.gcntZeroAt = c->gcntZeroAt;
  h.gstale = c->gstale;
  h.listsValid = c->listsValid;
  h.recFresh = c->recFresh;
  h.pvFresh = c->pvFresh;
  h.dFresh = c->dFresh;
  h.segsValid = c->segsValid;
  h.gridValid = c->gridValid;
  h.ssr = c->ssr;
  h.gs = c->gs;
  h.gc = c->gc;
  h.band = c->band;
  h.cur = c->cur;
  h.alt = c->alt;
  h.storeMem[0] = c->storeMem[0];
  h.storeMem[1] = c->storeMem[1];
  return h;
}
void loadHost(sphdem_ctx* c, const HostFields& h) {
  c->sphStep = h.sphStep;
  c->ng = h.ng;
  c->ns = h.ns;
  c->nd = h.nd;
  c->K = h.K;
  c->lgen = h.lgen;
  c->countsClean = h.countsClean;
  c->gcntZeroAt = h.gcntZeroAt;
  c->gstale = h.gstale;
  c->listsValid = h.listsValid;
  c->recFresh = h.recFresh;
  c->pvFresh = h.pvFresh;
  c->dFresh = h.dFresh;
  c->segsValid = h.segsValid;
  c->gridValid = h.gridValid;
  c->ssr = h.ssr;
  c->gs = h.gs;
  c->gc = h.gc;
  c->band = h.band;
  c->cur = h.cur;
  c->alt = h.alt;
  c->storeMem[0] = h.storeMem[0];
  c->storeMem[1] = h.storeMem[1];
}
// what the captured launches depend on: buffers, counts, grids, constants
void putRaw(std::vector<unsigned char>& v, const void* x, size_t bytes) {
  const unsigned char* b = static_cast<const unsigned char*>(x);
  v.insert(v.end(), b, b + bytes);
}
#define putBytes(v, x) putRaw(v, &(x), sizeof(x))
std::vector<unsigned char> graphKeyOf(const sphdem_ctx* c, int hookset) {
  std::vector<unsigned char> v;
  putBytes(v, c->allocEpoch);
  putBytes(v, c->storeMem[0]);
  putBytes(v, c->storeMem[1]);
  putBytes(v, c->cur.posh);
  const int ints[] = {c->n, c->ns, c->nd, c->ns0, c->nd0, c->K, c->gbase, hookset, c->tileSeg,
                      c->walkTPB, c->reachShep, c->overlap ? 1 : 0};
  putBytes(v, ints);
  putBytes(v, c->ghostCap);
  putBytes(v, c->ssr);
  putBytes(v, c->gs);
  putBytes(v, c->gc);
  putBytes(v, c->band);
  // field by field: these structs have padding
  putBytes(v, c->box.periodic);
  putBytes(v, c->box.Lp);
  putBytes(v, c->box.Lm);
  const CouplingRule& R = c->rule;
  const double rd[] = {R.cmin[0], R.cmin[1], R.cmin[2], R.e, R.s, R.thr, R.rc2, R.skip2,
                       R.gmin[0], R.gmin[1], R.gmin[2], R.gspace[0], R.gspace[1], R.gspace[2]};
  putBytes(v, rd);
  putBytes(v, R.nc);
  putBytes(v, c->P);
  return v;
}
#undef putBytes
std::atomic<long long> captureSerial{0};
void dropGraph(sphdem_ctx* c) {
  if (c->graphExec) (void)hipGraphExecDestroy(c->graphExec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graphExec = nullptr;
  c->graph = nullptr;
  c->graphKey.clear();
}
int captureStep(sphdem_ctx* ctx, int hookset) {
  dropGraph(ctx);
  if (!ctx->keepMem) {
    CHECK(allocStore(ctx, ctx->keep, ctx->keepMem, ctx->cap));
    HIPC(hipMalloc(&ctx->dgKeep, sizeof(DevGlobals)));
  }
  const bool periodic = ctx->box.periodic[0] || ctx->box.periodic[1] || ctx->box.periodic[2];
  ctx->graphG = periodic ? (int)std::min<int64_t>(ctx->ghostCap, (int64_t)ctx->ng + ctx->ng / 8 + 256)
                         : 0;
  const HostFields before = saveHost(ctx);
  HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->graphMode = true;
  hipLaunchKernelGGL(k_snapshot, dim3(blocks(ctx->n)), dim3(TPB), 0, ctx->stream, ctx->cur,
                     ctx->keep, ctx->n, ctx->dg, ctx->dgKeep, ctx->gate);
  int rc = hipGetLastError() == hipSuccess ? 0 : fail(ctx, SPHDEM_E_HIP, "k_snapshot");
  if (!rc) rc = runOneStep(ctx, hookset);
  ctx->graphMode = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  const HostFields after = saveHost(ctx);
  loadHost(ctx, before);
  if (rc || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : fail(ctx, SPHDEM_E_HIP, "graph capture: %s", hipGetErrorString(e));
  }
  hipGraphExec_t x = nullptr;
  if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess || !x) {
    (void)hipGraphDestroy(g);
    return fail(ctx, SPHDEM_E_HIP, "graph instantiation failed");
  }
  ctx->graph = g;
  ctx->graphExec = x;
  ctx->graphSerial = ++captureSerial;
  ctx->graphAfter = after;
  ctx->graphKey = graphKeyOf(ctx, hookset);
  return 0;
}
bool graphReady(const sphdem_ctx* ctx, int hookset) {
  return ctx->graphExec && ctx->graphKey == graphKeyOf(ctx, hookset);
}
// count replays on the context's stream, without waiting
int launchReplays(sphdem_ctx* ctx, int count) {
  HIPC(hipMemsetAsync(ctx->gate, 0, sizeof(int) * 4, ctx->stream));
  for (int r = 0; r < count; r++) HIPC(hipGraphLaunch(ctx->graphExec, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->gateHost, ctx->gate, sizeof(int) * 3, hipMemcpyDeviceToHost,
                      ctx->stream));
  HIPC(hipMemcpyAsync(ctx->gateHost + 3, &ctx->dg->ghostTotal, sizeof(int),
                      hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}
// after launchReplays: waits, then sets the host fields; *done = the replays
// that completed, *stopped = whether the gate stopped one (its step not run)
int finishReplays(sphdem_ctx* ctx, int count, int* done, bool* stopped) {
  HIPC(hipStreamSynchronize(ctx->stream));
  const long long step0 = ctx->sphStep;
  *stopped = ctx->gateHost[0] != 0;
  *done = *stopped ? ctx->gateHost[1] - 1 : count;
  loadHost(ctx, ctx->graphAfter);
  ctx->sphStep = step0 + *done;
  if (!*stopped) {
    if (ctx->gbase > 0) ctx->ng = ctx->gateHost[3];
    return 0;
  }
  // back to the state the stopped step started from.  The buffers a reset
  // rebuilds (cells, ghosts, lists) stay as the stopped step left them: the
  // caller runs that step as launches next, whose resets rebuild them before
  // anything reads them
  const size_t n = (size_t)ctx->n;
  for (int k = 0; k < kNumVec; k++)
    HIPC(hipMemcpyAsync(storeArr(ctx->cur, k), storeArr(ctx->keep, k), sizeof(double4) * n,
                        hipMemcpyDeviceToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->cur.iam, ctx->keep.iam, sizeof(int) * n, hipMemcpyDeviceToDevice,
                      ctx->stream));
  HIPC(hipMemcpyAsync(ctx->cur.orig, ctx->keep.orig, sizeof(int) * n, hipMemcpyDeviceToDevice,
                      ctx->stream));
  HIPC(hipMemcpyAsync(ctx->cur.cell, ctx->keep.cell, sizeof(int) * n, hipMemcpyDeviceToDevice,
                      ctx->stream));
  HIPC(hipMemcpyAsync(ctx->dg, ctx->dgKeep, sizeof(DevGlobals), hipMemcpyDeviceToDevice,
                      ctx->stream));
  ctx->listsValid = false;
  ctx->ng = 0;
  dropGraph(ctx);
  return 0;
}
// Steps launched one after another with the resets in their device form,
// nothing read back until finishReplays (the same gate, copy and recovery as
// the replays; without a capture, parity of the stores and reinitialisation
// steps need no special care).  The context must be in the state a reset
// leaves (any step run as launches).
bool deviceResetReady(const sphdem_ctx* c) {
  return c->n > 0 && c->gridValid && c->gcntZeroAt == c->n && c->countsClean >= c->K + 1 &&
         c->readyEpoch == c->allocEpoch;
}
// beginDeferred, then deferredStep per step (the steps of several contexts
// may interleave), then endDeferred, then finishReplays
int beginDeferred(sphdem_ctx* ctx) {
  if (!ctx->keepMem) {
    CHECK(allocStore(ctx, ctx->keep, ctx->keepMem, ctx->cap));
    HIPC(hipMalloc(&ctx->dgKeep, sizeof(DevGlobals)));
  }
  const bool periodic = ctx->box.periodic[0] || ctx->box.periodic[1] || ctx->box.periodic[2];
  ctx->graphG = periodic ? (int)std::min<int64_t>(ctx->ghostCap, (int64_t)ctx->ng + ctx->ng / 8 + 256)
                         : 0;
  HIPC(hipMemsetAsync(ctx->gate, 0, sizeof(int) * 4, ctx->stream));
  ctx->deferredStep0 = ctx->sphStep;
  return 0;
}
int deferredStep(sphdem_ctx* ctx, int hookset) {
  ctx->graphMode = true;
  hipLaunchKernelGGL(k_snapshot, dim3(blocks(ctx->n)), dim3(TPB), 0, ctx->stream, ctx->cur,
                     ctx->keep, ctx->n, ctx->dg, ctx->dgKeep, ctx->gate);
  int rc = hipGetLastError() == hipSuccess ? 0 : fail(ctx, SPHDEM_E_HIP, "k_snapshot");
  if (!rc) rc = runOneStep(ctx, hookset);
  ctx->graphMode = false;
  return rc;
}
int endDeferred(sphdem_ctx* ctx) {
  ctx->graphAfter = saveHost(ctx);
  ctx->sphStep = ctx->deferredStep0;  // finishReplays counts the steps that completed
  HIPC(hipMemcpyAsync(ctx->gateHost, ctx->gate, sizeof(int) * 3, hipMemcpyDeviceToHost,
                      ctx->stream));
  HIPC(hipMemcpyAsync(ctx->gateHost + 3, &ctx->dg->ghostTotal, sizeof(int),
                      hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

// One graph holding the captured step of every box of an ensemble as a child
// node, replayed with one launch per ensemble step on one stream.  It is kept
// for the next call that passes the same captures (each capture has a serial
// number no later capture repeats, so a box destroyed or captured again
// never matches) and rebuilt otherwise.
struct EnsembleGraph {
  std::vector<long long> key;  // the device, then each replaying box's capture serial
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr;
  int device = -1;
};
std::mutex ensembleMutex;
EnsembleGraph ensembleGraph;

// the replays of launchReplays for every box with count[b] > 0 (all the same
// count, all on one device), queued on the ensemble graph's stream; each
// box's gate copied back as launchReplays copies it.  The boxes' own streams
// are drained first, and finishReplays' wait on a box's stream is then
// preceded by the wait on the ensemble stream here.
int launchEnsemble(sphdem_ctx* const* ctxs, int nbox, const int* count, int replays) {
  std::lock_guard<std::mutex> lock(ensembleMutex);
  EnsembleGraph& E = ensembleGraph;
  const int device = ctxs[0]->device;
  std::vector<long long> key{device};
  for (int b = 0; b < nbox; b++)
    if (count[b]) key.push_back(ctxs[b]->graphSerial);
  sphdem_ctx* c0 = nullptr;
  for (int b = 0; b < nbox && !c0; b++)
    if (count[b]) c0 = ctxs[b];
  sphdem_ctx* const ctx = c0;  // where HIPC reports
  if (E.device != device || key != E.key) {
    if (E.exec) (void)hipGraphExecDestroy(E.exec);
    if (E.graph) (void)hipGraphDestroy(E.graph);
    E.exec = nullptr;
    E.graph = nullptr;
    E.key.clear();
    if (E.stream && E.device != device) {
      (void)hipSetDevice(E.device);
      (void)hipStreamDestroy(E.stream);
      E.stream = nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) return fail(c0, SPHDEM_E_HIP, "hipSetDevice");
    E.device = device;
    if (!E.stream && hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking) != hipSuccess)
      return fail(c0, SPHDEM_E_HIP, "ensemble stream");
    if (hipGraphCreate(&E.graph, 0) != hipSuccess) return fail(c0, SPHDEM_E_HIP, "hipGraphCreate");
    for (int b = 0; b < nbox; b++) {
      if (!count[b]) continue;
      hipGraphNode_t node;
      if (hipGraphAddChildGraphNode(&node, E.graph, nullptr, 0, ctxs[b]->graph) != hipSuccess)
        return fail(ctxs[b], SPHDEM_E_HIP, "ensemble graph: child node");
    }
    if (hipGraphInstantiate(&E.exec, E.graph, nullptr, nullptr, 0) != hipSuccess || !E.exec) {
      E.exec = nullptr;
      return fail(c0, SPHDEM_E_HIP, "ensemble graph instantiation failed");
    }
    E.key = key;
  }
  HIPC(hipSetDevice(device));
  for (int b = 0; b < nbox; b++) {
    if (!count[b]) continue;
    HIPC(hipStreamSynchronize(ctxs[b]->stream));
    HIPC(hipMemsetAsync(ctxs[b]->gate, 0, sizeof(int) * 4, E.stream));
  }
  for (int r = 0; r < replays; r++) HIPC(hipGraphLaunch(E.exec, E.stream));
  for (int b = 0; b < nbox; b++) {
    if (!count[b]) continue;
    sphdem_ctx* c = ctxs[b];
    HIPC(hipMemcpyAsync(c->gateHost, c->gate, sizeof(int) * 3, hipMemcpyDeviceToHost, E.stream));
    HIPC(hipMemcpyAsync(c->gateHost + 3, &c->dg->ghostTotal, sizeof(int), hipMemcpyDeviceToHost,
                        E.stream));
  }
  HIPC(hipStreamSynchronize(E.stream));
  return 0;
}

int graphSteps(sphdem_ctx* ctx, int nsteps, int hookset) {
  int done = 0;
  while (done < nsteps) {
    if (reinitDue(ctx, ctx->sphStep + 1)) {
      CHECK(runOneStep(ctx, hookset));
      done++;
      continue;
    }
    if (!graphReady(ctx, hookset)) {
      CHECK(runOneStep(ctx, hookset));  // the host step a capture follows
      done++;
      if (done < nsteps && !reinitDue(ctx, ctx->sphStep + 1)) CHECK(captureStep(ctx, hookset));
      continue;
    }
    int count = 0;
    while (done + count < nsteps && !reinitDue(ctx, ctx->sphStep + 1 + count)) count++;
    int ok = 0;
    bool stopped = false;
    CHECK(launchReplays(ctx, count));
    CHECK(finishReplays(ctx, count, &ok, &stopped));
    done += ok;
    if (stopped) {
      CHECK(runOneStep(ctx, hookset));
      done++;
    }
  }
  return 0;
}
}  // namespace

int sphdem_run_steps(sphdem_ctx* ctx, int nsteps, int hookset) {
  CHECK(runSteps(ctx, nsteps, hookset));
  return checkFlags(ctx);
}

int sphdem_run_steps_graph(sphdem_ctx* ctx, int nsteps, int hookset) {
  CHECK(checkLoaded(ctx));
  if (nsteps < 0 || hookset < SPHDEM_HOOKS_NONE || hookset > SPHDEM_HOOKS_SEDIMENT)
    return fail(ctx, SPHDEM_E_ARG, "bad nsteps/hookset");
  HIPC(hipSetDevice(ctx->device));
  // a timing window brackets launches one by one: those steps run as launches
  if (ctx->timedKernel >= 0) return sphdem_run_steps(ctx, nsteps, hookset);
  const bool ov = ctx->overlap;
  ctx->overlap = false;  // one stream: the captured step is one chain
  const int rc = graphSteps(ctx, nsteps, hookset);
  ctx->overlap = ov;
  CHECK(rc);
  return checkFlags(ctx);
}

namespace {
int manyGraph(sphdem_ctx* const* ctxs, int nbox, int nsteps, int hookset, int workers,
              int* failed_box, bool oneGraph) {
  if (failed_box) *failed_box = -1;
  if (nbox < 0 || nsteps < 0 || workers < 0 || (nbox > 0 && !ctxs)) return SPHDEM_E_ARG;
  if (hookset < SPHDEM_HOOKS_NONE || hookset > SPHDEM_HOOKS_SEDIMENT) return SPHDEM_E_ARG;
  if (nbox == 0 || nsteps == 0) return 0;
  std::vector<sphdem_ctx*> order(ctxs, ctxs + nbox);
  std::sort(order.begin(), order.end());
  if (!order[0] || std::adjacent_find(order.begin(), order.end()) != order.end())
    return SPHDEM_E_ARG;  // a null or a repeated context
  for (int b = 0; b < nbox; b++)
    if (!ctxs[b]->loaded) return fail(ctxs[b], SPHDEM_E_STATE, "no particles uploaded");
  if (workers == 0) workers = 8;
  workers = std::min(workers, nbox);
  std::vector<bool> overlap(nbox);
  for (int b = 0; b < nbox; b++) {
    overlap[b] = ctxs[b]->overlap;
    ctxs[b]->overlap = false;
  }
  std::atomic<int> err{0};
  std::atomic<int> errBox{-1};
  // workers take every workers-th box: a host step and a capture where a box
  // has no usable graph.  Then the replays: with oneGraph, and every box
  // replaying the same number of steps, one graph holding all the boxes'
  // captures is launched once per step on one stream; otherwise each box's
  // replays are queued on its own stream.  Then the waits (and, for a box
  // whose gate stopped a replay, the rest of its steps through graphSteps)
  std::vector<int> done(nbox, 0), count(nbox, 0);
  auto fault = [&](int b, int rc) {
    int none = 0;
    if (err.compare_exchange_strong(none, rc)) errBox.store(b);
  };
  auto parallel = [&](auto&& body) {
    std::vector<std::thread> pool;
    pool.reserve(workers - 1);
    auto run = [&](int w) {
      for (int b = w; b < nbox && !err.load(); b += workers) {
        const int rc = hipSetDevice(ctxs[b]->device) == hipSuccess ? body(b) : SPHDEM_E_HIP;
        if (rc) return fault(b, rc);
      }
    };
    for (int w = 1; w < workers; w++) pool.emplace_back(run, w);
    run(0);
    for (std::thread& t : pool) t.join();
  };
  parallel([&](int b) {
    sphdem_ctx* c = ctxs[b];
    bool plain = false;  // a reinit step in the window: this box runs step by step
    for (int k = 1; k <= nsteps && !plain; k++) plain = reinitDue(c, c->sphStep + k);
    int rc = 0;
    if (plain) {
      rc = graphSteps(c, nsteps, hookset);
      done[b] = nsteps;
    } else if (!graphReady(c, hookset)) {
      rc = runOneStep(c, hookset);
      done[b] = 1;
      if (!rc && nsteps > 1) rc = captureStep(c, hookset);
    }
    if (!rc) count[b] = nsteps - done[b];
    return rc;
  });
  int together = oneGraph ? -1 : 0;  // the replay count every replaying box shares
  for (int b = 0; b < nbox && together != 0; b++) {
    if (count[b] == 0) continue;
    if (together == -1) together = count[b];
    if (count[b] != together || ctxs[b]->device != ctxs[0]->device) together = 0;
  }
  if (!err.load() && together > 0) {
    const int rc = launchEnsemble(ctxs, nbox, count.data(), together);
    if (rc) fault(0, rc);
  } else if (!err.load()) {
    parallel([&](int b) { return count[b] ? launchReplays(ctxs[b], count[b]) : 0; });
  }
  if (!err.load())
    parallel([&](int b) {
      if (count[b] == 0) return 0;
      sphdem_ctx* c = ctxs[b];
      int ok = 0;
      bool stopped = false;
      int rc = finishReplays(c, count[b], &ok, &stopped);
      if (!rc && stopped) rc = graphSteps(c, nsteps - done[b] - ok, hookset);
      return rc;
    });
  for (int b = 0; b < nbox; b++) ctxs[b]->overlap = overlap[b];
  for (int b = 0; b < nbox && !err.load(); b++) {
    sphdem_ctx* ctx = ctxs[b];
    if (hipSetDevice(ctx->device) != hipSuccess) {
      err.store(SPHDEM_E_HIP);
      errBox.store(b);
      break;
    }
    const int rc = checkFlags(ctx);
    if (rc) {
      err.store(rc);
      errBox.store(b);
    }
  }
  if (failed_box) *failed_box = errBox.load();
  return err.load();
}
}  // namespace

int sphdem_run_steps_many_graph(sphdem_ctx* const* ctxs, int nbox, int nsteps, int hookset,
                                int workers, int* failed_box) {
  return manyGraph(ctxs, nbox, nsteps, hookset, workers, failed_box, false);
}

int sphdem_run_steps_ensemble_graph(sphdem_ctx* const* ctxs, int nbox, int nsteps, int hookset,
                                    int workers, int* failed_box) {
  return manyGraph(ctxs, nbox, nsteps, hookset, workers, failed_box, true);
}

int sphdem_run_steps_many(sphdem_ctx* const* ctxs, int nbox, int nsteps, int hookset,
                          int workers, int* failed_box) {
  if (failed_box) *failed_box = -1;
  if (nbox < 0 || nsteps < 0 || workers < 0 || (nbox > 0 && !ctxs)) return SPHDEM_E_ARG;
  if (nbox == 0 || nsteps == 0) return 0;
  std::vector<sphdem_ctx*> order(ctxs, ctxs + nbox);
  std::sort(order.begin(), order.end());
  if (!order[0] || std::adjacent_find(order.begin(), order.end()) != order.end())
    return SPHDEM_E_ARG;  // a null or a repeated context
  if (workers == 0) workers = 8;
  workers = std::min(workers, nbox);
  // the boxes overlap one another, so each box keeps its passes on one stream
  // (the extra fork/join of a second stream per box only costs here)
  std::vector<bool> overlap(nbox);
  for (int b = 0; b < nbox; b++) {
    overlap[b] = ctxs[b]->overlap;
    ctxs[b]->overlap = false;
  }
  std::atomic<int> err{0};
  std::atomic<int> errBox{-1};
  // each worker takes every workers-th box.  A box whose resets can run on
  // the device (any box that has run a step) gets its steps launched with
  // nothing read back, the worker launching one step of each of its boxes in
  // turn, and is waited for once at the end; a box that has not yet run a
  // step runs its first one as launches.  A box whose device reset met what
  // only the host handles is rolled back to the step it stopped at and
  // finishes on the host.
  auto work = [&](int w) {
    std::vector<int> mine, done, count;
    for (int b = w; b < nbox; b += workers) mine.push_back(b);
    done.assign(mine.size(), 0);
    count.assign(mine.size(), 0);
    auto fault = [&](int b, int rc) {
      int none = 0;
      if (err.compare_exchange_strong(none, rc)) errBox.store(b);
    };
    for (size_t i = 0; i < mine.size() && !err.load(); i++) {
      sphdem_ctx* c = ctxs[mine[i]];
      int rc = hipSetDevice(c->device) == hipSuccess ? 0 : SPHDEM_E_HIP;
      if (!rc) rc = checkLoaded(c);
      if (!rc && (hookset < SPHDEM_HOOKS_NONE || hookset > SPHDEM_HOOKS_SEDIMENT))
        rc = fail(c, SPHDEM_E_ARG, "bad nsteps/hookset");
      if (!rc && !deviceResetReady(c)) {
        rc = runOneStep(c, hookset);
        done[i] = 1;
      }
      if (!rc && done[i] < nsteps) {
        count[i] = nsteps - done[i];
        rc = beginDeferred(c);
      }
      if (rc) return fault(mine[i], rc);
    }
    for (int r = 0; r < nsteps && !err.load(); r++)
      for (size_t i = 0; i < mine.size() && !err.load(); i++) {
        if (r >= count[i]) continue;
        sphdem_ctx* c = ctxs[mine[i]];
        int rc = hipSetDevice(c->device) == hipSuccess ? 0 : SPHDEM_E_HIP;
        if (!rc) rc = deferredStep(c, hookset);
        if (!rc && r + 1 == count[i]) rc = endDeferred(c);
        if (rc) return fault(mine[i], rc);
      }
    for (size_t i = 0; i < mine.size() && !err.load(); i++) {
      if (count[i] == 0) continue;
      sphdem_ctx* c = ctxs[mine[i]];
      int ok = 0;
      bool stopped = false;
      int rc = hipSetDevice(c->device) == hipSuccess ? 0 : SPHDEM_E_HIP;
      if (!rc) rc = finishReplays(c, count[i], &ok, &stopped);
      if (!rc && stopped) rc = runSteps(c, count[i] - ok, hookset);
      if (rc) return fault(mine[i], rc);
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(workers - 1);
  for (int w = 1; w < workers; w++) pool.emplace_back(work, w);
  work(0);
  for (std::thread& t : pool) t.join();
  for (int b = 0; b < nbox; b++) ctxs[b]->overlap = overlap[b];
  // the last step's lists of every box (one round trip per box, after all ran)
  for (int b = 0; b < nbox && !err.load(); b++) {
    sphdem_ctx* ctx = ctxs[b];
    if (hipSetDevice(ctx->device) != hipSuccess) {
      err.store(SPHDEM_E_HIP);
      errBox.store(b);
      break;
    }
    const int rc = checkFlags(ctx);
    if (rc) {
      err.store(rc);
      errBox.store(b);
    }
  }
  if (failed_box) *failed_box = errBox.load();
  return err.load();
}

int sphdem_reset(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  return doReset(ctx);
}

int sphdem_get_cell_coords(sphdem_ctx* ctx, int32_t* ijk) {
  CHECK(checkLoaded(ctx));
  if (!ijk) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  int* buf = (int*)ctx->fieldbuf;
  hipLaunchKernelGGL(k_cell_coords, dim3(blocks(ctx->n)), dim3(TPB), 0, ctx->stream, ctx->cur,
                     ctx->n, ctx->gs, ctx->rule, buf);
  LAUNCHED();
  HIPC(hipMemcpyAsync(ijk, buf, sizeof(int) * 3 * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

// device: a, b, image are device pointers filled in place (no host copy)
static int getPairs(sphdem_ctx* ctx, int kind, int64_t* count, int32_t* a, int32_t* b,
                    int32_t* image, int64_t capacity, bool device) {
  CHECK(checkLoaded(ctx));
  if (!count || kind < 0 || kind > 3) return SPHDEM_E_ARG;
  if ((kind == SPHDEM_PAIRS_DEM || kind >= SPHDEM_PAIRS_SPH2DEM) && !liqDem(ctx)) {
    *count = 0;
    return 0;
  }
  HIPC(hipSetDevice(ctx->device));
  const int nq = (kind == SPHDEM_PAIRS_SPH || kind == SPHDEM_PAIRS_SPH2DEM) ? ctx->ns : ctx->nd;
  if (nq == 0) {
    *count = 0;
    return 0;
  }
  // the SPH set is read from the cached lists the SPH passes use (built now
  // if no pass has built them since the last reset)
  if (kind == SPHDEM_PAIRS_SPH) {
    CHECK(ensureLists(ctx));
    CHECK(checkFlags(ctx));
  }
  if (nq > ctx->pairCap) {
    if (ctx->pairCnt) HIPC(hipFree(ctx->pairCnt));
    if (ctx->pairOff) HIPC(hipFree(ctx->pairOff));
    ctx->pairCnt = nullptr;
    ctx->pairOff = nullptr;
    ctx->pairCap = 0;
    const int cap = nq + nq / 4;
    HIPC(hipMalloc(&ctx->pairCnt, sizeof(int) * cap));
    HIPC(hipMalloc(&ctx->pairOff, sizeof(long long) * cap));
    ctx->pairCap = cap;
  }
  int* cnt = ctx->pairCnt;
  long long* off = ctx->pairOff;
  HIPC(hipMemsetAsync(cnt, 0, sizeof(int) * nq, ctx->stream));
  auto enumerate = [&](int mode, const long long* o, int* da, int* db, int* di) -> int {
    if (kind == SPHDEM_PAIRS_SPH) {
      if (mode == 0)
        HIPC(hipMemcpyAsync(cnt, ctx->nbcnt, sizeof(int) * nq, hipMemcpyDeviceToDevice,
                            ctx->stream));
      else
        hipLaunchKernelGGL(k_list_pairs, dim3(blocks(nq)), dim3(TPB), 0, ctx->stream, ctx->cur,
                           ctx->ns, ctx->n, ctx->nbl, ctx->nbcnt, ctx->gcode, o, da, db, di);
      LAUNCHED();
      return 0;
    }
    hipLaunchKernelGGL(k_pairs, dim3(blocks(nq)), dim3(TPB), 0, ctx->stream, ctx->cur, ctx->ns,
                       ctx->nd, kind, ctx->gs, ctx->gc, ctx->start, ctx->box, ctx->rule,
                       ctx->reachFull, mode, cnt, o, da, db, di);
    LAUNCHED();
    return 0;
  };
  CHECK(enumerate(0, nullptr, nullptr, nullptr, nullptr));
  std::vector<int> hc(nq);
  HIPC(hipMemcpyAsync(hc.data(), cnt, sizeof(int) * nq, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  std::vector<long long> ho(nq);
  long long tot = 0;
  for (int i = 0; i < nq; i++) {
    ho[i] = tot;
    tot += hc[i];
  }
  *count = tot;
  if (a && b && image && tot > 0) {
    if (tot > capacity)
      return fail(ctx, SPHDEM_E_CAPACITY, "pair buffer of %lld < %lld", (long long)capacity, tot);
    int *da = nullptr, *db = nullptr, *di = nullptr;
    if (device) {
      da = a;
      db = b;
      di = image;
    } else {
      HIPC(hipMalloc(&da, sizeof(int) * tot));
      HIPC(hipMalloc(&db, sizeof(int) * tot));
      HIPC(hipMalloc(&di, sizeof(int) * tot));
    }
    HIPC(hipMemcpyAsync(off, ho.data(), sizeof(long long) * nq, hipMemcpyHostToDevice, ctx->stream));
    CHECK(enumerate(1, off, da, db, di));
    if (!device) {
      HIPC(hipMemcpyAsync(a, da, sizeof(int) * tot, hipMemcpyDeviceToHost, ctx->stream));
      HIPC(hipMemcpyAsync(b, db, sizeof(int) * tot, hipMemcpyDeviceToHost, ctx->stream));
      HIPC(hipMemcpyAsync(image, di, sizeof(int) * tot, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPC(hipStreamSynchronize(ctx->stream));
    if (!device) {
      (void)hipFree(da);
      (void)hipFree(db);
      (void)hipFree(di);
    }
  }
  return 0;
}

int sphdem_get_pairs(sphdem_ctx* ctx, int kind, int64_t* count, int32_t* a, int32_t* b,
                     int32_t* image, int64_t capacity) {
  return getPairs(ctx, kind, count, a, b, image, capacity, false);
}

int sphdem_get_pairs_device(sphdem_ctx* ctx, int kind, int64_t* count, int32_t* a, int32_t* b,
                            int32_t* image, int64_t capacity) {
  return getPairs(ctx, kind, count, a, b, image, capacity, true);
}

int sphdem_pass_omega(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  CHECK(passOmega(ctx));
  return checkFlags(ctx);
}
int sphdem_pass_eos(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  return passEos(ctx);
}
int sphdem_pass_force(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  CHECK(passForce(ctx));
  return checkFlags(ctx);
}
int sphdem_pass_dddt(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  CHECK(passDddt(ctx));
  return checkFlags(ctx);
}
int sphdem_pass_density(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  HIPC(hipSetDevice(ctx->device));
  CHECK(passDensity(ctx));
  return checkFlags(ctx);
}
int sphdem_pass_porosity(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  if (!liqDem(ctx)) return fail(ctx, SPHDEM_E_STATE, "LIQ_DEM is off");
  HIPC(hipSetDevice(ctx->device));
  return passCoupling(ctx);
}
int sphdem_pass_fpi(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  if (!liqDem(ctx)) return fail(ctx, SPHDEM_E_STATE, "LIQ_DEM is off");
  HIPC(hipSetDevice(ctx->device));
  return passFpi(ctx);
}
int sphdem_pass_dem_contacts(sphdem_ctx* ctx) {
  CHECK(checkLoaded(ctx));
  if (!liqDem(ctx)) return fail(ctx, SPHDEM_E_STATE, "LIQ_DEM is off");
  HIPC(hipSetDevice(ctx->device));
  return passContacts(ctx);
}
int sphdem_timing_begin(sphdem_ctx* ctx, const char* kernel) {
  if (!ctx || !kernel) return SPHDEM_E_ARG;
  for (int k = 0; k < KID_COUNT; k++)
    if (strcmp(kernel, kKernelNames[k]) == 0) {
      ctx->timedKernel = k;
      ctx->evUsed = 0;
      return 0;
    }
  return fail(ctx, SPHDEM_E_ARG, "unknown kernel name '%s'", kernel);
}

int sphdem_timing_end(sphdem_ctx* ctx, double* total_ms, int64_t* launches) {
  if (!ctx || !total_ms || !launches) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipStreamSynchronize(ctx->stream));
  double tot = 0.0;
  const size_t pairs = ctx->evUsed / 2;
  for (size_t i = 0; i < pairs; i++) {
    float ms = 0.0f;
    HIPC(hipEventElapsedTime(&ms, ctx->evPool[2 * i], ctx->evPool[2 * i + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = (int64_t)pairs;
  ctx->timedKernel = -1;
  ctx->evUsed = 0;
  return 0;
}

int sphdem_reduce_dt(sphdem_ctx* ctx, double* newdt) {
  CHECK(checkLoaded(ctx));
  if (!newdt) return SPHDEM_E_ARG;
  HIPC(hipSetDevice(ctx->device));
  DevGlobals d;
  HIPC(hipMemcpyAsync(&d, ctx->dg, sizeof(d), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  memcpy(newdt, &d.newDt_bits, sizeof(double));
  return 0;
}

}  // extern "C"
