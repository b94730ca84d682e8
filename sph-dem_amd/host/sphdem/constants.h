// constants.h -- builds the library's sphdem_params from the compile-time
// constants of a case's customConstants.h (the names the reference's core
// headers read: PERIODIC, RMIN, RMAX, DENS, REFD, GAMMA, HFAC, H, PSEP,
// SPSOUND, PRB, VISCOSITY, PI, the DEM_* set, COUPLING_RADIUS, ...) and its
// #defines (VAR_H_CORRECTION, LIQ_DEM, VISC_MORRIS, ...).  Include it after the
// case's customConstants.h.  The domain is the one splitDomain gives a single
// rank (misc.cpp:726-737): [RMIN, RMAX] on periodic axes, padded by a quarter
// of the box on the others.
#pragma once

#include <cmath>

#include "../../../include/sphdem_c.h"

namespace sphdem_host {

inline sphdem_params paramsFromConstants() {
  sphdem_params P = {};
  uint32_t f = 0;
#ifdef VAR_H_CORRECTION
  f |= SPHDEM_F_VAR_H_CORRECTION;
#endif
#ifdef CONST_H
  f |= SPHDEM_F_CONST_H;
#endif
#ifdef LIQ_DEM
  f |= SPHDEM_F_LIQ_DEM;
#endif
#ifdef VISC_MORRIS
  f |= SPHDEM_F_VISC_MORRIS;
#endif
#ifdef VISC_ARTIFICIAL
  f |= SPHDEM_F_VISC_ARTIFICIAL;
#endif
#ifdef LIQ_DEM_CUSTOM_WALL_CONTACTS
  f |= SPHDEM_F_CUSTOM_WALL;
#endif
#ifdef NO_DEM_CONTACTS
  f |= SPHDEM_F_NO_DEM_CONTACTS;
#endif
#ifdef FIXED_DEM
  f |= SPHDEM_F_FIXED_DEM;
#endif
#ifdef VISC_MONAGHAN
  f |= SPHDEM_F_VISC_MONAGHAN;
#endif
#ifdef VISC_CLEARY
  f |= SPHDEM_F_VISC_CLEARY;
#endif
#ifdef DEM_BLOCK
  f |= SPHDEM_F_DEM_BLOCK;
#endif
  P.flags = f;
  P.precision = SPHDEM_FP64;
  for (int i = 0; i < 3; i++) {
    P.periodic[i] = PERIODIC[i] ? 1 : 0;
    P.rmin[i] = RMIN[i];
    P.rmax[i] = RMAX[i];
    const double L = RMAX[i] - RMIN[i];
    P.proc_domain[2 * i] = PERIODIC[i] ? RMIN[i] : RMIN[i] - 0.25 * L;
    P.proc_domain[2 * i + 1] = PERIODIC[i] ? RMAX[i] : RMAX[i] + 0.25 * L;
  }
  P.reinit_every = REINIT_DENS_EVERY_N_STEPS;
  P.reinit_at = REINIT_DENS_AT_N_STEPS;
  P.damptime = DAMPTIME;
  P.pi = PI;
  P.dens = DENS;
  P.refd = REFD;
  P.gamma = GAMMA;
  P.hfac = HFAC;
  P.h = H;
  P.psep = PSEP;
  P.spsound = SPSOUND;
  P.prb = PRB;
  P.viscosity = VISCOSITY;
  P.alpha_artificial = ALPHA_ARTIFICIAL;
  P.wcon = WCON_WENDLAND;
  P.maxdt = MAXDT;
#ifdef LIQ_DEM
  P.dem_radius = DEM_RADIUS;
  P.dem_dens = DEM_DENS;
  P.dem_vol = DEM_VOL;
  P.dem_k = DEM_K;
  P.dem_gamma = DEM_GAMMA;
  P.dem_search_radius = DEM_SEARCH_RADIUS;
  P.coupling_radius = COUPLING_RADIUS;
  P.time_drop = TIME_DROP_PARTICLE;
  P.dem_timestep = DEM_TIMESTEP;
  P.liq_dem_timestep = LIQ_DEM_TIMESTEP;
  P.porosity_incr = POROSITY_INCR;
#endif
  P.drag_a = DRAG_A;  // Di Felice chi law, sph.h:1596
  P.drag_b = DRAG_B;
  return P;
}

}  // namespace sphdem_host
