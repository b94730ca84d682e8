// dataLL.h -- CdataLL for host code written against the reference's particle
// container API (dataLL.h:96-389): constructed over a particleContainer and a
// CglobalVars, reset(), traverse<F[, If]>() (with 0-3 extra arguments),
// neighbours<F[, If]>(), neighboursGroup<F[, If]>(), neighboursGroupCoupling
// <F, If>() with host functors, getParticles(), numParticles() and the public
// `globals`.  The particles live on the GPU in the library's own layout
// (libsphdem.so, include/sphdem_c.h); the container is a host view that is
// refreshed before a host functor runs and written back after it.  The pair
// traversals walk the library's own neighbour sets (sphdem_get_pairs: the
// cached lists of the last reset, or a fresh coupling search), in its
// cell-sorted query order and the reference's list order; a partner across a
// periodic boundary is handed over as a ghost copy shifted by the box length,
// as the reference's procGhostParticles are, and writes to ghost copies are
// dropped (the reference keeps them only until its next ghost sync).  The
// library's timestep passes (CsphIncompress) do not go through these: they run
// on the GPU.
//
// Errors behave as in the reference: a message on cerr, then exit(-1)
// (dataLL.cpp:200-211).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <deque>
#include <iostream>
#include <vector>

#include "../../../include/sphdem_c.h"
#include "globalVars.h"
#include "particle.h"

namespace sphdem_host {

class CdataLL {
 public:
  // CdataLL(particleContainer&, CglobalVars&, bool opt) (dataLL.cpp:8-40).  The
  // physics constants a case's customConstants.h fixes at compile time come in
  // as `params` (see constants.h); opt (spatial reordering) is implied: the
  // library keeps the particles cell-sorted.
  CdataLL(particleContainer& in_ps, CglobalVars& in_globals, bool opt, const sphdem_params& params,
          int device = 0)
      : globals(in_globals), ps(in_ps), params(params) {
    (void)opt;
    const int64_t n = (int64_t)ps.size();
    if (sphdem_create(&params, device, n > 0 ? n : 1, &ctx) != SPHDEM_OK) {
      std::cerr << "CdataLL: no usable gfx950 GPU (sphdem_create failed)" << std::endl;
      std::exit(-1);
    }
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
      ps[i].tag = (int)i + 1;
    }
    check(sphdem_upload(ctx, n, r.data(), v.data(), vh.data(), m.data(), h.data(), d.data(),
                        iam.data()));
    pushGlobals();
  }
  ~CdataLL() {
    if (ctx) sphdem_destroy(ctx);
  }
  CdataLL(const CdataLL&) = delete;
  CdataLL& operator=(const CdataLL&) = delete;

  // reset() (dataLL.cpp:140-236): wrap, delete out-of-range particles, rebin.
  void reset() { check(sphdem_reset(ctx)); }

  // traverse<F>() / traverse<F, If>() (dataLL.impl.h:16-120)
  template <void Thefunct(Cparticle&, CglobalVars&)>
  void traverse() {
    pull();
    for (Cparticle& p : ps)
      if (p.iam != deleted) Thefunct(p, globals);
    push();
  }
  template <void Thefunct(Cparticle&, CglobalVars&), bool Iffunct(Cparticle&)>
  void traverse() {
    pull();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals);
    push();
  }
  template <void Thefunct(Cparticle&, CglobalVars&, double), bool Iffunct(Cparticle&)>
  void traverse(double a) {
    pull();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals, a);
    push();
  }

  template <void Thefunct(Cparticle&, CglobalVars&, double)>
  void traverse(double a) {
    traverse<Thefunct, always>(a);
  }
  template <class T, void Thefunct(Cparticle&, CglobalVars&, T&), bool Iffunct(Cparticle&)>
  void traverse(T& param) {
    pull();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals, param);
    push();
  }
  template <void Thefunct(Cparticle&, CglobalVars&, double, double), bool Iffunct(Cparticle&)>
  void traverse(double a, double b) {
    pull();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals, a, b);
    push();
  }
  template <void Thefunct(Cparticle&, CglobalVars&, double, double, double),
            bool Iffunct(Cparticle&)>
  void traverse(double a, double b, double c) {
    pull();
    for (Cparticle& p : ps)
      if (p.iam != deleted && Iffunct(p)) Thefunct(p, globals, a, b, c);
    push();
  }

  // neighbours<F[, If]>() (dataLL.impl.h:165-204): F(pa, pb, globals) for every
  // neighbour pb of every particle pa (that passes If), over the cached
  // neighbour sets of the last reset.
  template <void Thefunct(Cparticle&, Cparticle&, CglobalVars&)>
  void neighbours() {
    neighbours<Thefunct, always>();
  }
  template <void Thefunct(Cparticle&, Cparticle&, CglobalVars&), bool Iffunct(Cparticle&)>
  void neighbours() {
    pull();
    Groups G = groups(SPHDEM_PAIRS_SPH, SPHDEM_PAIRS_DEM);
    for (Group& q : G.list)
      if (Iffunct(ps[q.a]))
        for (Cparticle* pb : q.nb) Thefunct(ps[q.a], *pb, globals);
    push();
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
    pull();
    Groups G = groups(SPHDEM_PAIRS_SPH, SPHDEM_PAIRS_DEM);
    for (Group& q : G.list)
      if (Iffunct(ps[q.a])) Thefunct(ps[q.a], q.nb, globals);
    push();
  }
  template <class T, void Thefunct(Cparticle&, std::vector<Cparticle*>&, CglobalVars&, T&),
            bool Iffunct(Cparticle&)>
  void neighboursGroup(T& param) {
    pull();
    Groups G = groups(SPHDEM_PAIRS_SPH, SPHDEM_PAIRS_DEM);
    for (Group& q : G.list)
      if (Iffunct(ps[q.a])) Thefunct(ps[q.a], q.nb, globals, param);
    push();
  }
  // neighboursGroupCoupling<F, If>() (dataLL.impl.h:256-267): the coupling
  // search (calcNeighboursCoupling, dataLL.cpp:1104-1192) done afresh: DEM
  // partners of a fluid particle, fluid partners of a DEM particle.
  template <void Thefunct(Cparticle&, std::vector<Cparticle*>&, CglobalVars&),
            bool Iffunct(Cparticle&)>
  void neighboursGroupCoupling() {
    pull();
    Groups G = groups(SPHDEM_PAIRS_SPH2DEM, SPHDEM_PAIRS_DEM2SPH);
    for (Group& q : G.list)
      if (Iffunct(ps[q.a])) Thefunct(ps[q.a], q.nb, globals);
    push();
  }

  int numParticles() const { return (int)ps.size(); }
  // getParticles() (dataLL.h:238): the host view, refreshed from the GPU.
  particleContainer* getParticles() {
    pull();
    return &ps;
  }

  // Host view <-> GPU.  pull() copies every field and the globals to the host;
  // push() writes every settable field and the globals back.
  void pull() {
    pullGlobals();
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
  }
  void push() {
    const size_t n = ps.size();
    std::vector<double> b(3 * n);
    for (const VecField& f : vecFields()) {
      for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) b[3 * i + k] = (ps[i].*f.m)[k];
      check(sphdem_set_field(ctx, f.id, b.data()));
    }
    for (const SclField& f : sclFields()) {
      for (size_t i = 0; i < n; i++) b[i] = ps[i].*f.m;
      check(sphdem_set_field(ctx, f.id, b.data()));
    }
    pushGlobals();
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

  // Device forms of the common CcustomSim traversals (no host round trip).
  void deviceBodyForce(uint32_t typeMask, double ax, double ay, double az) {
    const double a[3] = {ax, ay, az};
    check(sphdem_hook_body_force(ctx, typeMask, a));
  }
  void deviceDemFloor() { check(sphdem_hook_dem_floor(ctx)); }
  void deviceFreeze(uint32_t typeMask, double tFreeze) {
    check(sphdem_hook_freeze(ctx, typeMask, tFreeze));
  }
  void deviceSumDrag() { check(sphdem_hook_sum_drag(ctx)); }

  sphdem_ctx* context() { return ctx; }
  void check(int rc) {
    if (rc != SPHDEM_OK) {
      std::cerr << "sphdem: " << sphdem_last_error(ctx) << " (code " << rc << ")" << std::endl;
      std::exit(-1);
    }
  }

  CglobalVars& globals;

 private:
  static bool always(Cparticle&) { return true; }

  // The neighbour sets of a traversal: one group per live particle of the
  // two query kinds (fluid / DEM), in the library's query order, then the
  // particles without partners by index; ghost copies of periodic partners
  // live in `ghosts` (a deque, so pointers to them stay valid).
  struct Group {
    int a;
    std::vector<Cparticle*> nb;
  };
  struct Groups {
    std::vector<Group> list;
    std::deque<Cparticle> ghosts;
  };
  Groups groups(int kindFluid, int kindDem) {
    Groups G;
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
        if (img[k] != 13) {  // a periodic image: a shifted copy of b
          G.ghosts.push_back(ps[b[k]]);
          Cparticle& g = G.ghosts.back();
          const int s[3] = {img[k] % 3 - 1, (img[k] / 3) % 3 - 1, img[k] / 9 - 1};
          for (int d = 0; d < 3; d++)
            if (s[d] != 0) g.r[d] = (s[d] > 0 ? params.rmax[d] - params.rmin[d]
                                               : params.rmin[d] - params.rmax[d]) + g.r[d];
          pb = &g;
        }
        G.list[slot[a[k]]].nb.push_back(pb);
      }
    }
    for (int i = 0; i < n; i++)
      if (slot[i] < 0 && ps[i].iam != deleted) G.list.push_back(Group{i, {}});
    return G;
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
};

}  // namespace sphdem_host
