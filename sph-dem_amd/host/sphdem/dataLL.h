// dataLL.h -- CdataLL for host code written against the reference's particle
// container API (dataLL.h:96-270): constructed over a particleContainer and a
// CglobalVars as the reference's is (the physics constants come from the
// case's customConstants.h, as the reference's headers take them), reset(),
// traverse<F[, If]>() with 0-3 extra arguments, neighbours<F[, If]>(),
// neighboursGroup<F[, If]>(), neighboursGroupCoupling<F, If>(),
// neighboursUsing<F>(), neighboursGroupAtRadius<F, If>(), calcNeighbours(),
// sync/reverseSyncParticlesBetweenProcs, setGlobalTimestep, calcTraverseOrder,
// insertNewParticle / markForDeletion / deleteParticles, getParticleFromTag,
// sumOverProcs, getParticles(), numParticles() and the public `globals`.
//
// The particles live on the GPU in the library's own layout (libsphdem.so,
// include/sphdem_c.h); the container is a host view kept in sync lazily: it
// is downloaded only when the device has changed since the last download,
// and host writes are found by comparing each field with its last synced
// value, so only the fields a functor changed are uploaded, right before the
// next call that reads the device state.  The pair traversals walk the
// library's own neighbour sets (sphdem_get_pairs: the cached lists of the last
// reset, or a fresh coupling search), in its cell-sorted query order and the
// reference's list order; a partner across a periodic boundary is handed over
// as a ghost copy shifted by the box length, as the reference's
// procGhostParticles are.  The ghost copies of the last pair traversal stay
// until the next one: reverseSyncParticlesBetweenProcs carries their values
// back to their origins, syncParticlesBetweenProcs refreshes them.  The
// library's timestep passes (CsphIncompress) do not go through these: they run
// on the GPU.
//
// Errors behave as in the reference: a message on cerr, then exit(-1)
// (dataLL.cpp:200-211, 241-244).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <iostream>
#include <map>
#include <vector>

#include "../../../include/sphdem_c.h"
#include "constants.h"
#include "customConstants.h"
#include "globalVars.h"
#include "particle.h"
#include "sphConstants.h"

namespace sphdem_host {

class CdataLL {
 public:
  // CdataLL(particleContainer&, CglobalVars&, bool opt) (dataLL.h:103,
  // dataLL.cpp:8-40), with the constants of the case's customConstants.h;
  // opt (spatial reordering) is implied: the library keeps the particles
  // cell-sorted.  The device holds up to MAX_NUM_PARTICLES_PER_CPU particles.
  CdataLL(particleContainer& in_ps, CglobalVars& in_globals, bool opt)
      : CdataLL(in_ps, in_globals, opt, paramsFromConstants()) {}
  // the same with the constants given at run time (an ensemble member, a
  // calibration run) instead of the compiled ones
  CdataLL(particleContainer& in_ps, CglobalVars& in_globals, bool opt, const sphdem_params& params,
          int device = 0)
      : globals(in_globals), ps(in_ps), params(params) {
    (void)opt;
    const int64_t n = (int64_t)ps.size();
    capacity = std::max<int64_t>({n, (int64_t)MAX_NUM_PARTICLES_PER_CPU, 1});
    if (sphdem_create(&params, device, capacity, &ctx) != SPHDEM_OK) {
      std::cerr << "CdataLL: no usable gfx950 GPU (sphdem_create failed)" << std::endl;
      std::exit(-1);
    }
    for (int64_t i = 0; i < n; i++) ps[i].tag = (int)i + 1;
    uploadAll(false);
  }
  ~CdataLL() {
    if (ctx) sphdem_destroy(ctx);
  }
  CdataLL(const CdataLL&) = delete;
  CdataLL& operator=(const CdataLL&) = delete;

  // reset() (dataLL.cpp:140-236): wrap, delete out-of-range particles, rebin.
  void reset() {
    flush();
    check(sphdem_reset(ctx));
    deviceChanged();
  }

  // ---- traverse<F>() / traverse<F, If>() (dataLL.impl.h:16-163): every live
  // particle of the container (ghosts are not handed to host functors)
  template <void Thefunct(Cparticle&, CglobalVars&)>
  void traverse() {
    traverse<Thefunct, always>();
  }
  template <void Thefunct(Cparticle&, CglobalVars&), bool Iffunct(Cparticle&)>
  void traverse() {
    begin();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals);
    end();
  }
  template <void Thefunct(Cparticle&, CglobalVars&, double)>
  void traverse(double a) {
    traverse<Thefunct, always>(a);
  }
  template <void Thefunct(Cparticle&, CglobalVars&, double), bool Iffunct(Cparticle&)>
  void traverse(double a) {
    begin();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals, a);
    end();
  }
  template <class T, void Thefunct(Cparticle&, CglobalVars&, T&)>
  void traverse(T& param) {
    traverse<T, Thefunct, always>(param);
  }
  template <class T, void Thefunct(Cparticle&, CglobalVars&, T&), bool Iffunct(Cparticle&)>
  void traverse(T& param) {
    begin();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals, param);
    end();
  }
  template <void Thefunct(Cparticle&, CglobalVars&, double, double)>
  void traverse(double a, double b) {
    traverse<Thefunct, always>(a, b);
  }
  template <void Thefunct(Cparticle&, CglobalVars&, double, double), bool Iffunct(Cparticle&)>
  void traverse(double a, double b) {
    begin();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals, a, b);
    end();
  }
  template <void Thefunct(Cparticle&, CglobalVars&, double, double, double),
            bool Iffunct(Cparticle&)>
  void traverse(double a, double b, double c) {
    begin();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals, a, b, c);
    end();
  }

  // ---- pair traversals over the library's neighbour sets
  // neighbours<F[, If]>() (dataLL.impl.h:165-204): F(pa, pb, globals) for every
  // neighbour pb of every particle pa (that passes If), over the cached
  // neighbour sets of the last reset.
  template <void Thefunct(Cparticle&, Cparticle&, CglobalVars&)>
  void neighbours() {
    neighbours<Thefunct, always>();
  }
  template <void Thefunct(Cparticle&, Cparticle&, CglobalVars&), bool Iffunct(Cparticle&)>
  void neighbours() {
    begin();
    groups(SPHDEM_PAIRS_SPH, SPHDEM_PAIRS_DEM);
    for (Group& q : G.list)
      if (Iffunct(ps[q.a]))
        for (Cparticle* pb : q.nb) Thefunct(ps[q.a], *pb, globals);
    end();
  }
  // neighboursGroup<F[, If]>() (dataLL.impl.h:206-240): F(pa, neighbours of pa,
  // globals) for every particle pa that passes If.
  template <void Thefunct(Cparticle&, std::vector<Cparticle*>&, CglobalVars&)>
  void neighboursGroup() {
    neighboursGroup<Thefunct, always>();
  }
  template <void Thefunct(Cparticle&, std::vector<Cparticle*>&, CglobalVars&),
            bool Iffunct(Cparticle&)>
  void neighboursGroup() {
    begin();
    groups(SPHDEM_PAIRS_SPH, SPHDEM_PAIRS_DEM);
    for (Group& q : G.list)
      if (Iffunct(ps[q.a])) Thefunct(ps[q.a], q.nb, globals);
    end();
  }
  template <class T, void Thefunct(Cparticle&, std::vector<Cparticle*>&, CglobalVars&, T&),
            bool Iffunct(Cparticle&)>
  void neighboursGroup(T& param) {
    begin();
    groups(SPHDEM_PAIRS_SPH, SPHDEM_PAIRS_DEM);
    for (Group& q : G.list)
      if (Iffunct(ps[q.a])) Thefunct(ps[q.a], q.nb, globals, param);
    end();
  }
  // neighboursGroupCoupling<F, If>() (dataLL.impl.h:256-267): the coupling
  // search (calcNeighboursCoupling, dataLL.cpp:1104-1192) done afresh: DEM
  // partners of a fluid particle, fluid partners of a DEM particle.
  template <void Thefunct(Cparticle&, std::vector<Cparticle*>&, CglobalVars&),
            bool Iffunct(Cparticle&)>
  void neighboursGroupCoupling() {
    begin();
    groups(SPHDEM_PAIRS_SPH2DEM, SPHDEM_PAIRS_DEM2SPH);
    for (Group& q : G.list)
      if (Iffunct(ps[q.a])) Thefunct(ps[q.a], q.nb, globals);
    end();
  }

  // ---- searches on the host view (helpers of analysis and output code)
  // calcNeighbours (dataLL.cpp:1026-1075) of any particle, e.g. one outside
  // the container: the container's particles b (and periodic images) with
  // len2(r_p - r_b) <= (2 max(h_p, h_b))^2, DEM among DEM and the rest among
  // sph/sphBoundary, in x, y, z cell order of a grid of edge 2 hmax, then
  // container order within a cell
  void calcNeighbours(std::vector<Cparticle*>& out, Cparticle& p) {
    begin();
    hostGrid();
    const bool isDem = p.iam == dem;
    hostSearch(p, 1, [&](Cparticle* b) {
      if ((b->iam == dem) != isDem) return;
      const double lim = 2.0 * std::max(p.h, b->h);
      if (len2(p.r - b->r) <= lim * lim) out.push_back(b);
    });
  }
  // calcNeighboursAtRadius (dataLL.cpp:1078-1100): sph/sphBoundary particles
  // other than p with len2(r_p - r_b) < radius^2
  void calcNeighboursAtRadius(std::vector<Cparticle*>& out, Cparticle& p, const double radius,
                              const double radius2) {
    begin();
    hostGrid();
    const int reach = std::max(1, (int)std::ceil(radius / hg.edge));
    hostSearch(p, reach, [&](Cparticle* b) {
      if (b->iam == dem || b == &p) return;
      if (len2(p.r - b->r) < radius2) out.push_back(b);
    });
  }
  // neighboursUsing<F>(ps) (dataLL.impl.h:283-306): F over each given particle
  // and its neighbours in the container
  template <void Thefunct(Cparticle&, Cparticle&, CglobalVars&)>
  void neighboursUsing(std::vector<Cparticle>& others) {
    for (Cparticle& q : others) {
      std::vector<Cparticle*> nb;
      calcNeighbours(nb, q);
      for (Cparticle* b : nb) Thefunct(q, *b, globals);
    }
    end();
  }
  template <void Thefunct(Cparticle&, std::vector<Cparticle*>&, CglobalVars&)>
  void neighboursUsing(std::vector<Cparticle>& others) {
    for (Cparticle& q : others) {
      std::vector<Cparticle*> nb;
      calcNeighbours(nb, q);
      Thefunct(q, nb, globals);
    }
    end();
  }
  // neighboursGroupAtRadius<F, If>(radius) (dataLL.impl.h:243-254)
  template <void Thefunct(Cparticle&, std::vector<Cparticle*>&, CglobalVars&),
            bool Iffunct(Cparticle&)>
  void neighboursGroupAtRadius(const double radius) {
    begin();
    for (Cparticle& p : ps) {
      if (p.iam == deleted || !Iffunct(p)) continue;
      std::vector<Cparticle*> nb;
      calcNeighboursAtRadius(nb, p, radius, radius * radius);
      Thefunct(p, nb, globals);
    }
    end();
  }

  // ---- ghosts of the last pair traversal (dataLL.impl.h:348-477)
  // syncParticlesBetweenProcs<T, read, write>(): each ghost takes its
  // origin's value
  template <class T, T readFunct(Cparticle&), void writeFunct(Cparticle&, T)>
  void syncParticlesBetweenProcs() {
    for (size_t g = 0; g < G.ghosts.size(); g++)
      writeFunct(G.ghosts[g], readFunct(ps[G.ghostOrigin[g]]));
  }
  // reverseSyncParticlesBetweenProcs<T, read, write>(): each ghost's value is
  // handed to writeFunct on its origin (a sum for the scatter operators), in
  // the order the ghosts were made
  template <class T, T readFunct(Cparticle&), void writeFunct(Cparticle&, T)>
  void reverseSyncParticlesBetweenProcs() {
    begin();
    for (size_t g = 0; g < G.ghosts.size(); g++)
      writeFunct(ps[G.ghostOrigin[g]], readFunct(G.ghosts[g]));
    end();
  }

  // setGlobalTimestep (dataLL.cpp:774-782): the minimum over ranks, one rank
  void setGlobalTimestep(double dt) { globals.dt = dt; }
  // sumOverProcs (dataLL.cpp:790-832): the identity at one rank
  void sumOverProcs(vect*, int) {}
  void sumOverProcs(double*, int) {}
  // calcTraverseOrder (dataLL.cpp:935-952) reorders the traversal along a
  // Morton curve for cache locality; the library keeps the particles in cell
  // order on the device at every reset, so there is nothing to do
  void calcTraverseOrder() {}
  // updateParticles (dataLL.cpp:852-855): rebuild everything from the
  // container after wholesale changes on the host
  void updateParticles() {
    structureChanged = true;
    flush();
  }
  void printMe() {
    int64_t n = 0, del = 0;
    sphdem_num_particles(ctx, &n);
    sphdem_num_deleted(ctx, &del);
    std::cout << "CdataLL: " << n << " particles on the GPU (" << del
              << " deleted as out of range), capacity " << capacity << std::endl;
  }

  // ---- insertion and deletion (dataLL.cpp:238-313)
  void insertNewParticle(Cparticle& newP) {
    begin();
    ps.push_back(newP);
    if ((int64_t)ps.size() > capacity) {
      std::cout << "insertNewParticle(): number of particles greater than  max limit!!!"
                << std::endl;
      std::exit(-1);
    }
    structureChanged = true;
    tagIndex.clear();
  }
  void markForDeletion(Cparticle& p) { p.tag = -111; }
  // every particle tagged -111 leaves the container; the last particle fills
  // each hole, as deleteParticle does
  void deleteParticles() {
    begin();
    for (size_t i = 0; i < ps.size();) {
      if (ps[i].tag == -111) {
        ps[i] = ps.back();
        ps.pop_back();
        structureChanged = true;
        continue;
      }
      i++;
    }
    tagIndex.clear();
  }
  // getParticleFromTag (dataLL.cpp:1195-1214): NULL for an unknown tag
  Cparticle* getParticleFromTag(int tag) {
    begin();
    if (tagIndex.empty())
      for (size_t i = 0; i < ps.size(); i++) tagIndex[ps[i].tag] = i;
    auto it = tagIndex.find(tag);
    if (it == tagIndex.end() || it->second >= ps.size() || ps[it->second].tag != tag)
      return nullptr;
    return &ps[it->second];
  }

  int numParticles() const { return (int)ps.size(); }
  // getParticles() (dataLL.h:238): the host view, current with the GPU; the
  // caller's writes reach the GPU before the next device call
  particleContainer* getParticles() {
    begin();
    return &ps;
  }

  // ---- host view <-> GPU
  // begin(): the host view holds the device state (downloaded if the device
  // changed since); every host write after it is found by flush().  The
  // globals belong to the host between device calls: flush() hands them to
  // the device, deviceChanged() takes them back.
  void begin() {
    if (hostCurrent) {
      hostWrites = true;
      return;
    }
    flushStructure();
    const size_t n = ps.size();
    std::vector<double> b(3 * n);
    for (const VecField& f : vecFields()) {
      check(sphdem_get_field(ctx, f.id, b.data()));
      for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) (ps[i].*f.m)[k] = b[3 * i + k];
    }
    for (const SclField& f : sclFields()) {
      check(sphdem_get_field(ctx, f.id, b.data()));
      for (size_t i = 0; i < n; i++) ps[i].*f.m = b[i];
    }
    std::vector<int32_t> t(n);
    check(sphdem_get_field(ctx, SPHDEM_FIELD_IAM, t.data()));
    for (size_t i = 0; i < n; i++) ps[i].iam = t[i];
    snap = ps;
    hostCurrent = true;
    hostWrites = true;
  }
  // end of a host traversal: the writes stay on the host until flush()
  void end() {}
  // flush(): the fields the host changed since the last sync go to the GPU
  // (compared field by field with their synced values), before any call that
  // reads the device state
  void flush() {
    if (structureChanged) {
      flushStructure();
    } else if (hostWrites) {
      const size_t n = ps.size();
      std::vector<double> b(3 * n);
      for (const VecField& f : vecFields()) {
        bool changed = false;
        for (size_t i = 0; i < n && !changed; i++)
          for (int k = 0; k < 3; k++)
            if (std::memcmp(&(ps[i].*f.m)[k], &(snap[i].*f.m)[k], sizeof(double)) != 0)
              changed = true;
        if (!changed) continue;
        for (size_t i = 0; i < n; i++)
          for (int k = 0; k < 3; k++) b[3 * i + k] = (ps[i].*f.m)[k];
        check(sphdem_set_field(ctx, f.id, b.data()));
        for (size_t i = 0; i < n; i++) snap[i].*f.m = ps[i].*f.m;
        uploads++;
      }
      for (const SclField& f : sclFields()) {
        bool changed = false;
        for (size_t i = 0; i < n && !changed; i++)
          if (std::memcmp(&(ps[i].*f.m), &(snap[i].*f.m), sizeof(double)) != 0) changed = true;
        if (!changed) continue;
        for (size_t i = 0; i < n; i++) b[i] = ps[i].*f.m;
        check(sphdem_set_field(ctx, f.id, b.data()));
        for (size_t i = 0; i < n; i++) snap[i].*f.m = ps[i].*f.m;
        uploads++;
      }
      hostWrites = false;
    }
    pushGlobals();
  }
  // the device state changed: the host view is stale
  void deviceChanged() {
    hostCurrent = false;
    hostWrites = false;
    pullGlobals();
  }
  // full download / full upload, for callers that manage the view themselves
  void pull() {
    flush();
    hostCurrent = false;
    begin();
  }
  void push() {
    structureChanged = true;
    flush();
  }
  void pullGlobals() {
    sphdem_globals g;
    check(sphdem_get_globals(ctx, &g));
    globals.time = g.time;
    globals.dt = g.dt;
    globals.newDt = g.newDt;
    globals.maxdt = g.maxdt;
    globals.sphStep = g.sphStep;
    for (int i = 0; i < 9; i++) globals.custom[i] = g.custom[i];
  }
  void pushGlobals() {
    sphdem_globals g;
    g.time = globals.time;
    g.dt = globals.dt;
    g.newDt = globals.newDt;
    g.maxdt = globals.maxdt;
    g.sphStep = globals.sphStep;
    for (int i = 0; i < 9; i++) g.custom[i] = globals.custom[i];
    check(sphdem_set_globals(ctx, &g));
  }
  // field uploads made by flush() so far (a measure of the lazy sync)
  long fieldUploads() const { return uploads; }

  // Device forms of the common CcustomSim traversals (no host round trip).
  void deviceBodyForce(uint32_t typeMask, double ax, double ay, double az) {
    const double a[3] = {ax, ay, az};
    flush();
    check(sphdem_hook_body_force(ctx, typeMask, a));
    deviceChanged();
  }
  void deviceDemFloor() {
    flush();
    check(sphdem_hook_dem_floor(ctx));
    deviceChanged();
  }
  void deviceFreeze(uint32_t typeMask, double tFreeze) {
    flush();
    check(sphdem_hook_freeze(ctx, typeMask, tFreeze));
    deviceChanged();
  }
  void deviceSumDrag() {
    flush();
    check(sphdem_hook_sum_drag(ctx));
    deviceChanged();
  }

  sphdem_ctx* context() { return ctx; }
  const sphdem_params& parameters() const { return params; }
  void check(int rc) {
    if (rc != SPHDEM_OK) {
      std::cerr << "sphdem: " << sphdem_last_error(ctx) << " (code " << rc << ")" << std::endl;
      std::exit(-1);
    }
  }

  CglobalVars& globals;

 private:
  static bool always(Cparticle&) { return true; }

  // re-uploads the whole container (after insertions or deletions): the
  // library's upload, then every field and the globals
  void flushStructure() {
    if (structureChanged) uploadAll(true);
  }
  void uploadAll(bool allFields) {
    const int64_t n = (int64_t)ps.size();
    std::vector<double> r(3 * n), v(3 * n), vh(3 * n), m(n), h(n), d(n);
    std::vector<int32_t> iam(n);
    for (int64_t i = 0; i < n; i++) {
      const Cparticle& p = ps[i];
      for (int k = 0; k < 3; k++) {
        r[3 * i + k] = p.r[k];
        v[3 * i + k] = p.v[k];
        vh[3 * i + k] = p.vhat[k];
      }
      m[i] = p.mass;
      h[i] = p.h;
      d[i] = p.dens;
      iam[i] = p.iam;
    }
    check(sphdem_upload(ctx, n, r.data(), v.data(), vh.data(), m.data(), h.data(), d.data(),
                        iam.data()));
    if (allFields) {
      std::vector<double> b(3 * n);
      for (const VecField& f : vecFields()) {
        for (int64_t i = 0; i < n; i++)
          for (int k = 0; k < 3; k++) b[3 * i + k] = (ps[i].*f.m)[k];
        check(sphdem_set_field(ctx, f.id, b.data()));
      }
      for (const SclField& f : sclFields()) {
        for (int64_t i = 0; i < n; i++) b[i] = ps[i].*f.m;
        check(sphdem_set_field(ctx, f.id, b.data()));
      }
    }
    pushGlobals();
    structureChanged = false;
    hostCurrent = false;
    hostWrites = false;
    snap.clear();
    G = Groups();
  }

  // The neighbour sets of a traversal: one group per live particle of the
  // two query kinds (fluid / DEM), in the library's query order, then the
  // particles without partners by index; ghost copies of periodic partners
  // live in `ghosts` (a deque, so pointers to them stay valid) with their
  // origins.
  struct Group {
    int a;
    std::vector<Cparticle*> nb;
  };
  struct Groups {
    std::vector<Group> list;
    std::deque<Cparticle> ghosts;
    std::vector<int> ghostOrigin;
  };
  void groups(int kindFluid, int kindDem) {
    flush();
    G = Groups();
    const int n = (int)ps.size();
    std::vector<int> slot(n, -1);
    for (int kind : {kindFluid, kindDem}) {
      int64_t cnt = 0;
      check(sphdem_get_pairs(ctx, kind, &cnt, nullptr, nullptr, nullptr, 0));
      if (cnt == 0) continue;
      std::vector<int32_t> a(cnt), b(cnt), img(cnt);
      check(sphdem_get_pairs(ctx, kind, &cnt, a.data(), b.data(), img.data(), cnt));
      for (int64_t k = 0; k < cnt; k++) {
        if (slot[a[k]] < 0) {
          slot[a[k]] = (int)G.list.size();
          G.list.push_back(Group{a[k], {}});
        }
        Cparticle* pb = &ps[b[k]];
        if (img[k] != 13) pb = image(b[k], img[k]);
        G.list[slot[a[k]]].nb.push_back(pb);
      }
    }
    for (int i = 0; i < n; i++)
      if (slot[i] < 0 && ps[i].iam != deleted) G.list.push_back(Group{i, {}});
  }
  // a ghost copy of particle b shifted by the periodic image code
  // (sx+1) + 3(sy+1) + 9(sz+1), with the reference's arithmetic
  // (r + (RMAX-RMIN) or r + (RMIN-RMAX) per shifted axis, dataLL.cpp:469-484)
  Cparticle* image(int b, int code) {
    G.ghosts.push_back(ps[b]);
    G.ghostOrigin.push_back(b);
    Cparticle& g = G.ghosts.back();
    const int s[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
    for (int d = 0; d < 3; d++)
      if (s[d] != 0)
        g.r[d] = (s[d] > 0 ? params.rmax[d] - params.rmin[d] : params.rmin[d] - params.rmax[d]) +
                 g.r[d];
    return &g;
  }

  // host grid of the searches on the host view: cells of edge 2 hmax over
  // the live particles and their periodic images within 2 hmax of a face
  struct HostEntry {
    Cparticle* p;
    int origin;  // index of the real particle (for images)
  };
  struct HostGrid {
    double edge = 0.0, lo[3] = {0, 0, 0};
    int dim[3] = {1, 1, 1};
    std::vector<std::vector<HostEntry>> cells;
    std::deque<Cparticle> images;
  } hg;
  void hostGrid() {
    double hmax = 0.0, lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (Cparticle& p : ps) {
      if (p.iam == deleted) continue;
      hmax = std::max(hmax, p.h);
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], p.r[d]);
        hi[d] = std::max(hi[d], p.r[d]);
      }
    }
    hg = HostGrid();
    hg.edge = std::max(2.0 * hmax, 1e-300);
    const double L[3] = {params.rmax[0] - params.rmin[0], params.rmax[1] - params.rmin[1],
                         params.rmax[2] - params.rmin[2]};
    std::vector<std::pair<Cparticle*, int>> all;
    for (size_t i = 0; i < ps.size(); i++)
      if (ps[i].iam != deleted) all.push_back({&ps[i], (int)i});
    // periodic images of the particles within one cell edge of a face
    for (size_t i = 0; i < ps.size(); i++) {
      if (ps[i].iam == deleted) continue;
      for (int code = 0; code < 27; code++) {
        if (code == 13) continue;
        const int s[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
        bool ok = true;
        for (int d = 0; d < 3 && ok; d++) {
          if (s[d] == 0) continue;
          if (!params.periodic[d]) ok = false;
          else if (s[d] > 0 && ps[i].r[d] > params.rmin[d] + hg.edge) ok = false;
          else if (s[d] < 0 && ps[i].r[d] < params.rmax[d] - hg.edge) ok = false;
        }
        if (!ok) continue;
        hg.images.push_back(ps[i]);
        Cparticle& g = hg.images.back();
        for (int d = 0; d < 3; d++)
          if (s[d] != 0) g.r[d] = (s[d] > 0 ? L[d] : -L[d]) + g.r[d];
        all.push_back({&g, (int)i});
      }
    }
    for (int d = 0; d < 3; d++) {
      if (params.periodic[d]) {
        lo[d] = std::min(lo[d], params.rmin[d] - hg.edge);
        hi[d] = std::max(hi[d], params.rmax[d] + hg.edge);
      }
      hg.lo[d] = lo[d] - hg.edge;
      hg.dim[d] = std::max(1, (int)((hi[d] - hg.lo[d]) / hg.edge) + 2);
    }
    hg.cells.assign((size_t)hg.dim[0] * hg.dim[1] * hg.dim[2], {});
    for (auto& e : all) {
      int c[3];
      cellOf(e.first->r, c);
      hg.cells[((size_t)c[0] * hg.dim[1] + c[1]) * hg.dim[2] + c[2]].push_back(
          HostEntry{e.first, e.second});
    }
  }
  void cellOf(const vect& r, int* c) const {
    for (int d = 0; d < 3; d++)
      c[d] = std::min(std::max((int)((r[d] - hg.lo[d]) / hg.edge), 0), hg.dim[d] - 1);
  }
  template <class Fn>
  void hostSearch(Cparticle& p, int reach, Fn fn) {
    int c[3];
    cellOf(p.r, c);
    for (int x = std::max(c[0] - reach, 0); x <= std::min(c[0] + reach, hg.dim[0] - 1); x++)
      for (int y = std::max(c[1] - reach, 0); y <= std::min(c[1] + reach, hg.dim[1] - 1); y++)
        for (int z = std::max(c[2] - reach, 0); z <= std::min(c[2] + reach, hg.dim[2] - 1); z++)
          for (HostEntry& e : hg.cells[((size_t)x * hg.dim[1] + y) * hg.dim[2] + z])
            fn(e.p);
  }

  struct VecField {
    int id;
    vect Cparticle::*m;
  };
  struct SclField {
    int id;
    double Cparticle::*m;
  };
  static const std::vector<VecField>& vecFields() {
    static const std::vector<VecField> f = {
        {SPHDEM_FIELD_R, &Cparticle::r},         {SPHDEM_FIELD_V, &Cparticle::v},
        {SPHDEM_FIELD_V0, &Cparticle::v0},       {SPHDEM_FIELD_VHAT, &Cparticle::vhat},
        {SPHDEM_FIELD_VHAT0, &Cparticle::vhat0}, {SPHDEM_FIELD_F, &Cparticle::f},
        {SPHDEM_FIELD_FP, &Cparticle::fp},       {SPHDEM_FIELD_FV, &Cparticle::fv},
        {SPHDEM_FIELD_FB, &Cparticle::fb},       {SPHDEM_FIELD_FDRAG, &Cparticle::fdrag},
        {SPHDEM_FIELD_FVEL, &Cparticle::fvel}};
    return f;
  }
  static const std::vector<SclField>& sclFields() {
    static const std::vector<SclField> f = {
        {SPHDEM_FIELD_DENS, &Cparticle::dens},         {SPHDEM_FIELD_DDDT, &Cparticle::dddt},
        {SPHDEM_FIELD_H, &Cparticle::h},               {SPHDEM_FIELD_MASS, &Cparticle::mass},
        {SPHDEM_FIELD_PRESS, &Cparticle::press},       {SPHDEM_FIELD_PDR2, &Cparticle::pdr2},
        {SPHDEM_FIELD_OMEGA, &Cparticle::omega},       {SPHDEM_FIELD_POROSITY, &Cparticle::porosity},
        {SPHDEM_FIELD_SHEPSUM, &Cparticle::shepSum},   {SPHDEM_FIELD_U, &Cparticle::u},
        {SPHDEM_FIELD_DUDT, &Cparticle::dudt}};
    return f;
  }

  particleContainer& ps;
  sphdem_params params;
  sphdem_ctx* ctx = nullptr;
  int64_t capacity = 0;
  bool hostCurrent = false, hostWrites = false, structureChanged = false;
  particleContainer snap;  // the host view as last synced with the GPU
  Groups G;                // the last pair traversal's groups and ghosts
  std::map<int, size_t> tagIndex;
  long uploads = 0;
};

}  // namespace sphdem_host

using sphdem_host::CdataLL;
