/*
 * sphdem_adjoint.h -- C ABI of the pair passes of the differentiable SPH-DEM
 * step: forward and hand-written backward (vector-Jacobian product) kernels of
 * the SPH force, the DEM contact and the fluid-particle drag passes.
 *
 * None of these has a counterpart in the reference, which has no derivative
 * code: they exist for gradient-based calibration of the drag closure, the
 * contact stiffness and the fluid viscosity.  Each forward kernel restates one
 * reference operator over an explicit pair set:
 *   sphdem_adj_force      CsphIncompress::calcForce (sphIncompress.h:899-972)
 *                         with calcViscForce (sph.h:579-666, VISC_MORRIS and
 *                         LIQ_DEM) and calcPressForce (sph.h:692-734)
 *   sphdem_adj_fpi        calcFPI (sph.h:1708-1792) with the Di Felice branch
 *                         of calcParticleDrag (sph.h:1575-1631)
 *   sphdem_adj_contacts   calcDEMContacts (sph.h:1676-1706) with the LINEAR
 *                         demNormal (sph.h:1495-1516)
 * and each backward kernel returns the adjoints of every differentiable input
 * given the adjoints of the outputs.
 *
 * Conventions
 *   - All arrays are device memory, indexed by original particle index;
 *     vectors are n*3 doubles (x,y,z), scalars n doubles.
 *   - Pairs are (a, b, image) triples of sphdem_get_pairs: b's periodic image
 *     sits at r_b + ((img%3)-1, (img/3)%3-1, img/9-1) * shift_len.
 *   - Forward outputs and backward adjoints are ADDED to the arrays given (the
 *     caller zeroes them); the sums use atomics, so the last bits vary run to
 *     run.
 *   - Every call is ordered on `stream` (a hipStream_t; NULL = the null
 *     stream) and returns 0, or SPHDEM_E_ARG for a null pointer or a negative
 *     count, or SPHDEM_E_HIP if a launch fails.
 */
#ifndef SPHDEM_ADJOINT_H
#define SPHDEM_ADJOINT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constants the pair passes read (the customConstants.h values of the case). */
typedef struct sphdem_adj_consts {
  double wcon;         /* WCON_WENDLAND                                   */
  double viscosity;    /* VISCOSITY                                       */
  double dens;         /* DENS (reference fluid density)                  */
  double pi;           /* PI                                              */
  double dem_radius;   /* DEM_RADIUS                                      */
  double dem_vol;      /* DEM_VOL                                         */
  double dem_k;        /* DEM_K                                           */
  double dem_gamma;    /* DEM_GAMMA                                       */
  double drag_a;       /* 3.7 of the Di Felice exponent                   */
  double drag_b;       /* 0.65 of the Di Felice exponent                  */
  double shift_len[3]; /* periodic box lengths (0 on a non-periodic axis) */
} sphdem_adj_consts;

/* Per-particle inputs.  Passes read only the members they name below. */
typedef struct sphdem_adj_fields {
  const double* r;     /* n*3 */
  const double* vhat;  /* n*3 */
  const double* h;
  const double* dens;
  const double* mass;
  const double* omega;
  const double* pdr2;
  const double* fpfb;  /* n*3: fp + fb of the fluid (fpi) */
  const double* fv;    /* n*3: viscous force of the fluid (fpi) */
  const double* por;   /* fluid porosity (fpi) */
  const int32_t* iam;  /* particle type (force: viscous term on SPH only) */
} sphdem_adj_fields;

/* Adjoint outputs, accumulated.  params[] receives d/d of
 * {viscosity, drag_a, drag_b, dem_k, dem_gamma}. */
typedef struct sphdem_adj_grads {
  double* r;
  double* vhat;
  double* h;
  double* dens;
  double* omega;
  double* pdr2;
  double* fpfb;
  double* fv;
  double* por;
  double* params; /* 5 doubles */
} sphdem_adj_grads;

/* SPH force over the SPH pair set: fv[a] += viscous, fp[a] += pressure.
 * Reads r, vhat, h, dens, mass, omega, pdr2, iam. */
int sphdem_adj_force(int64_t npairs, const int32_t* a, const int32_t* b, const int32_t* img,
                     const sphdem_adj_fields* in, const sphdem_adj_consts* c, double* fv,
                     double* fp, void* stream);
/* Backward of sphdem_adj_force given gfv, gfp (n*3): adds to grads r, vhat,
 * h, dens, omega, pdr2 and params[0]. */
int sphdem_adj_force_bwd(int64_t npairs, const int32_t* a, const int32_t* b, const int32_t* img,
                         const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                         const double* gfv, const double* gfp, const sphdem_adj_grads* g,
                         void* stream);

/* Fluid-particle interaction and drag of each DEM particle q = query[i]
 * over its fluid partners b[off[i] .. off[i+1]) (the DEM->SPH coupling set):
 * fdrag[q] = the calcFPI drag (zero when the Shepard sum is <= 0.5).
 * Reads r, vhat, h, dens, mass, fpfb, fv, por. */
int sphdem_adj_fpi(int32_t nquery, const int32_t* query, const int64_t* off, const int32_t* b,
                   const int32_t* img, const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                   double* fdrag, void* stream);
/* Backward of sphdem_adj_fpi given gfdrag (n*3): adds to grads r, vhat, h,
 * dens, fpfb, fv, por and params[0..2]. */
int sphdem_adj_fpi_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                       const int32_t* b, const int32_t* img, const sphdem_adj_fields* in,
                       const sphdem_adj_consts* c, const double* gfdrag,
                       const sphdem_adj_grads* g, void* stream);

/* DEM contacts over the DEM pair set: fc[a] += normal force / mass.
 * Reads r, vhat, mass. */
int sphdem_adj_contacts(int64_t npairs, const int32_t* a, const int32_t* b, const int32_t* img,
                        const sphdem_adj_fields* in, const sphdem_adj_consts* c, double* fc,
                        void* stream);
/* Backward of sphdem_adj_contacts given gfc (n*3): adds to grads r, vhat and
 * params[3..4]. */
int sphdem_adj_contacts_bwd(int64_t npairs, const int32_t* a, const int32_t* b,
                            const int32_t* img, const sphdem_adj_fields* in,
                            const sphdem_adj_consts* c, const double* gfc,
                            const sphdem_adj_grads* g, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SPHDEM_ADJOINT_H */
