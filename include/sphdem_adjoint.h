/*
 * sphdem_adjoint.h -- C ABI of the pair passes of the differentiable SPH-DEM
 * step: forward and hand-written backward (vector-Jacobian product) kernels of
 * the SPH force, omega, dddt, the fluid-side coupling sums, the fluid-particle
 * drag and the DEM contact passes.
 *
 * None of these has a counterpart in the reference, which has no derivative
 * code: they exist for gradient-based calibration of the drag closure, the
 * contact stiffness and the fluid viscosity.  Each forward kernel restates one
 * reference operator over an explicit pair set:
 *   sphdem_adj_force      CsphIncompress::calcForce (sphIncompress.h:899-972)
 *                         with calcViscForce (sph.h:579-666, VISC_MORRIS and
 *                         LIQ_DEM) and calcPressForce (sph.h:692-734)
 *   sphdem_adj_omega      Nsph::calcOmega (sph.h:437-452)
 *   sphdem_adj_dddt       CsphIncompress::calcDddtDudt (sphIncompress.h:783-851)
 *   sphdem_adj_coupling   calcPorosity, calcShepSum and calcDEMtoSPHDrag
 *                         (sph.h:1303-1362, with calcPorosityIncr 1246-1256)
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
 *     sits at r_b + ((img%3)-1, (img/3)%3-1, img/9-1) * shift_len.  The
 *     per-query passes take them grouped by a: query[i] = a, its partners
 *     b[off[i] .. off[i+1]) with their images; one thread per query keeps a's
 *     sums in registers.
 *   - Forward outputs and backward adjoints are ADDED to the arrays given (the
 *     caller zeroes them).  The backward kernels of force, omega and dddt
 *     gather: the thread of a particle also adds the terms of the pairs in
 *     which it is the partner, which requires a symmetric pair set (every
 *     (a, b, img) has its (b, a, 26 - img)); the SPH set of sphdem_get_pairs
 *     is symmetric (the criterion uses max(h_a, h_b)).  The coupling, fpi
 *     and contact backward kernels add partner terms with FP64 atomics, so
 *     their last bits vary run to run.
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
  double coupling_radius; /* LIQ_DEM_COUPLING_RADIUS                      */
  double porosity_incr;   /* (4/3) PI DEM_RADIUS^3                        */
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
  const double* fdrag; /* n*3: DEM drag of the last step (coupling) */
  const double* shep;  /* DEM Shepard sums of the forward (coupling backward) */
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
  double* fdrag;
  double* params; /* 5 doubles */
} sphdem_adj_grads;

/* SPH force of each query over its SPH pairs: fv[a] += viscous, fp[a] +=
 * pressure.  Reads r, vhat, h, dens, mass, omega, pdr2, iam. */
int sphdem_adj_force(int32_t nquery, const int32_t* query, const int64_t* off, const int32_t* b,
                     const int32_t* img, const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                     double* fv, double* fp, void* stream);
/* Backward of sphdem_adj_force given gfv, gfp (n*3): adds to grads r, vhat,
 * h, dens, omega, pdr2 and params[0]. */
int sphdem_adj_force_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                         const int32_t* b, const int32_t* img, const sphdem_adj_fields* in,
                         const sphdem_adj_consts* c, const double* gfv, const double* gfp,
                         const sphdem_adj_grads* g, void* stream);

/* omega[a] += h_a S_a / (3 rho_a), S_a = m_a dW/dh(0, h_a) + sum_b m_b
 * dW/dh(r_ab, h_a) (the caller starts omega at 1).  Reads r, h, dens, mass. */
int sphdem_adj_omega(int32_t nquery, const int32_t* query, const int64_t* off, const int32_t* b,
                     const int32_t* img, const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                     double* omega, void* stream);
/* Backward of sphdem_adj_omega given gomega (n): adds to grads r, h, dens. */
int sphdem_adj_omega_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                         const int32_t* b, const int32_t* img, const sphdem_adj_fields* in,
                         const sphdem_adj_consts* c, const double* gomega,
                         const sphdem_adj_grads* g, void* stream);

/* dddt[a] += sum_b m_b (vhat_a - vhat_b).gradW(r_ab, h_a)/omega_a.
 * Reads r, vhat, h, mass, omega. */
int sphdem_adj_dddt(int32_t nquery, const int32_t* query, const int64_t* off, const int32_t* b,
                    const int32_t* img, const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                    double* dddt, void* stream);
/* Backward of sphdem_adj_dddt given gdddt (n): adds to grads r, vhat, h,
 * omega. */
int sphdem_adj_dddt_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                        const int32_t* b, const int32_t* img, const sphdem_adj_fields* in,
                        const sphdem_adj_consts* c, const double* gdddt,
                        const sphdem_adj_grads* g, void* stream);

/* Fluid-side coupling of each fluid query over its DEM partners d (the
 * SPH->DEM coupling set), W = W(r_ad, Rc/2):
 *   por_sum[a] += POROSITY_INCR * sum_d W,   shep[d] += (m_a/rho_a) W,
 *   drag[a] += -sum_d m_d W fdrag_d / (rho_a shep_d)   (shep_d > 0)
 * (two launches: the drag needs the complete Shepard sums).  Reads r, dens,
 * mass, fdrag. */
int sphdem_adj_coupling(int32_t nquery, const int32_t* query, const int64_t* off,
                        const int32_t* d, const int32_t* img, const sphdem_adj_fields* in,
                        const sphdem_adj_consts* c, double* por_sum, double* shep, double* drag,
                        void* stream);
/* Backward of sphdem_adj_coupling given gpor (n) and gdrag (n*3), with in->shep
 * the forward's Shepard sums and work n zeroed doubles (the Shepard-sum
 * adjoints): adds to grads r, dens, fdrag. */
int sphdem_adj_coupling_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                            const int32_t* d, const int32_t* img, const sphdem_adj_fields* in,
                            const sphdem_adj_consts* c, const double* gpor, const double* gdrag,
                            double* work, const sphdem_adj_grads* g, void* stream);

/* sphdem_adj_coupling_bwd without the work argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * Shepard-sum adjoints go to a scratch array the library holds in device
 * memory, one per GPU for the life of the process (8 MiB: at most
 * SPHDEM_ADJ_NOALLOC_MAX_PARTICLES doubles); n is the particle count, which
 * must exceed every index in query and d.  The call clears scratch[0, n) with
 * a kernel, then runs the same two kernels as sphdem_adj_coupling_bwd; it
 * calls no hipMalloc, hipFree or allocating library routine.  Returns
 * SPHDEM_E_CAPACITY when n exceeds the maximum.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other host
 * threads) race and give wrong adjoints; calls ordered on one stream do not. */
#define SPHDEM_ADJ_NOALLOC_MAX_PARTICLES (1 << 20)
int sphdem_adj_coupling_bwd_noalloc(int32_t nquery, const int32_t* query, const int64_t* off,
                                    const int32_t* d, const int32_t* img,
                                    const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                                    const double* gpor, const double* gdrag, int64_t n,
                                    const sphdem_adj_grads* g, void* stream);

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
