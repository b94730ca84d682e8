/*
 * sphdem_c.h -- C ABI of the MI355X (gfx950) SPH-DEM per-timestep particle loop.
 *
 * The library replaces the per-timestep loop of SPH-DEM: the CdataLL particle
 * container (dataLL.h:96-389), the Nsph pair/group operators (sph.h) and the
 * CsphIncompress step driver (sphIncompress.cpp:3-251).  All state lives on one
 * GPU; every call is ordered on the context's HIP stream.
 *
 * Conventions
 *   - Every entry point returns int: 0 = ok, < 0 = SPHDEM_E_*.  The library never
 *     calls exit(); sphdem_last_error() describes the last failure.
 *   - Particles are addressed by their index 0..n-1 in the order they were
 *     uploaded (the "original index"; the reference's tag is index+1).  The
 *     device keeps them in its own cell-sorted order; get/set calls translate.
 *   - Vector fields are exchanged as n*3 doubles (x,y,z per particle), scalar
 *     fields as n doubles, integer fields as n int32.
 *   - One context per stream; calls on one context must not overlap.
 */
#ifndef SPHDEM_C_H
#define SPHDEM_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------- */
#define SPHDEM_OK 0
#define SPHDEM_E_ARG (-1)      /* bad argument (null pointer, size, field id)   */
#define SPHDEM_E_HIP (-2)      /* HIP runtime failure                           */
#define SPHDEM_E_STATE (-3)    /* call out of order (e.g. no particles loaded)  */
#define SPHDEM_E_CAPACITY (-4) /* grid or particle capacity exceeded             */
#define SPHDEM_E_DEVICE (-5)   /* no usable gfx950 device                        */

/* ---- particle types: the integer values of enum iamTypes (particle.h:17) --- */
#define SPHDEM_SPH 0
#define SPHDEM_SPH_BOUNDARY 1
#define SPHDEM_IMMOVABLE 2
#define SPHDEM_BOUNDARY 3
#define SPHDEM_GHOST 4
#define SPHDEM_DEM 11
#define SPHDEM_DELETED 255 /* no counterpart: marks a particle removed as out of range */

/* ---- compile-time physics switches of customConstants.h, as runtime bits --- */
#define SPHDEM_F_VAR_H_CORRECTION (1u << 0)  /* VAR_H_CORRECTION                 */
#define SPHDEM_F_CONST_H (1u << 1)           /* CONST_H                          */
#define SPHDEM_F_LIQ_DEM (1u << 2)           /* LIQ_DEM                          */
#define SPHDEM_F_VISC_MORRIS (1u << 3)       /* VISC_MORRIS                      */
#define SPHDEM_F_VISC_ARTIFICIAL (1u << 4)   /* VISC_ARTIFICIAL                  */
#define SPHDEM_F_CUSTOM_WALL (1u << 5)       /* LIQ_DEM_CUSTOM_WALL_CONTACTS     */
#define SPHDEM_F_NO_DEM_CONTACTS (1u << 6)   /* NO_DEM_CONTACTS                  */
#define SPHDEM_F_FIXED_DEM (1u << 7)         /* FIXED_DEM                        */
#define SPHDEM_F_VISC_MONAGHAN (1u << 8)     /* VISC_MONAGHAN                    */
#define SPHDEM_F_VISC_CLEARY (1u << 9)       /* VISC_CLEARY                      */
#define SPHDEM_F_DEM_BLOCK (1u << 10)        /* DEM_BLOCK                        */
/* At most one of VISC_MORRIS, VISC_CLEARY and VISC_MONAGHAN (calcViscForce,
 * sph.h:579-631, declares its `visc` once per variant, so the reference does
 * not compile with two); sphdem_create / sphdem_set_params refuse more. */

/* Precision of the pair arithmetic. */
#define SPHDEM_FP64 0
#define SPHDEM_MIXED 1

/*
 * Every customConstants.h value the loop reads (compile-time constants in the
 * reference; Appendix of constants: NDIM=3, WENDLAND kernel, KERNAL_RADIUS=2).
 * Derived constants are passed explicitly so that every implementation uses the
 * same double for them.
 */
typedef struct sphdem_params {
  uint32_t flags;          /* SPHDEM_F_* */
  int32_t precision;       /* SPHDEM_FP64 | SPHDEM_MIXED */
  int32_t periodic[3];     /* PERIODIC[] */
  int32_t reinit_every;    /* REINIT_DENS_EVERY_N_STEPS */
  int32_t reinit_at;       /* REINIT_DENS_AT_N_STEPS */
  int32_t damptime;        /* DAMPTIME (an int in the reference) */
  double rmin[3], rmax[3]; /* RMIN[], RMAX[] */
  double proc_domain[6];   /* globals.procDomain: xmin,xmax,ymin,ymax,zmin,zmax */
  double pi;               /* PI (3.14159265 in the reference cases) */
  double dens;             /* DENS */
  double refd;             /* REFD */
  double gamma;            /* GAMMA (EOS exponent under LIQ_DEM) */
  double hfac;             /* HFAC */
  double h;                /* H: initial hmax / search radius */
  double psep;             /* PSEP */
  double spsound;          /* SPSOUND */
  double prb;              /* PRB */
  double viscosity;        /* VISCOSITY (kinematic) */
  double alpha_artificial; /* ALPHA_ARTIFICIAL */
  double wcon;             /* WCON_WENDLAND = 21/(256 PI) */
  double maxdt;            /* globals.maxdt = MAXTIME/OUTSTEP */
  double dem_radius;       /* DEM_RADIUS */
  double dem_dens;         /* DEM_DENS */
  double dem_vol;          /* DEM_VOL */
  double dem_k;            /* DEM_K */
  double dem_gamma;        /* DEM_GAMMA */
  double dem_search_radius;/* DEM_SEARCH_RADIUS (= DEM_RADIUS for LINEAR) */
  double coupling_radius;  /* LIQ_DEM_COUPLING_RADIUS */
  double time_drop;        /* TIME_DROP_PARTICLE */
  double dem_timestep;     /* DEM_TIMESTEP (demConstants.h:6) */
  double liq_dem_timestep; /* LIQ_DEM_TIMESTEP (demConstants.h:19) */
  double porosity_incr;    /* (4/3) PI DEM_RADIUS^3, the factor of calcPorosityIncr */
  double drag_a;           /* 3.7  of the Di Felice exponent chi = a - b exp(..) */
  double drag_b;           /* 0.65 of the Di Felice exponent */
} sphdem_params;

/* The CglobalVars members the loop reads and writes (globalVars.h:11-235). */
typedef struct sphdem_globals {
  double time;
  double dt;
  double newDt;
  double maxdt;
  int64_t sphStep;
  double custom[9];
} sphdem_globals;

/* Field ids for get/set.  Widths: vectors 3, scalars 1, iam/tag int32. */
enum sphdem_field {
  SPHDEM_FIELD_R = 0,
  SPHDEM_FIELD_V,
  SPHDEM_FIELD_V0,
  SPHDEM_FIELD_VHAT,
  SPHDEM_FIELD_VHAT0,
  SPHDEM_FIELD_F,
  SPHDEM_FIELD_FP,
  SPHDEM_FIELD_FV,
  SPHDEM_FIELD_FB,
  SPHDEM_FIELD_FDRAG,
  SPHDEM_FIELD_FVEL,
  SPHDEM_FIELD_DENS = 16,
  SPHDEM_FIELD_DDDT,
  SPHDEM_FIELD_H,
  SPHDEM_FIELD_MASS,
  SPHDEM_FIELD_PRESS,
  SPHDEM_FIELD_PDR2,
  SPHDEM_FIELD_OMEGA,
  SPHDEM_FIELD_POROSITY,
  SPHDEM_FIELD_SHEPSUM,
  SPHDEM_FIELD_U,
  SPHDEM_FIELD_DUDT,
  SPHDEM_FIELD_IAM = 32,
  SPHDEM_FIELD_TAG
};

/* Pair-set kinds for sphdem_get_pairs (parity of the neighbour searches). */
enum sphdem_pair_kind {
  SPHDEM_PAIRS_SPH = 0,      /* calcNeighbours, SPH branch (dataLL.cpp:1049-1071)   */
  SPHDEM_PAIRS_DEM = 1,      /* calcNeighbours, DEM branch (dataLL.cpp:1028-1047)   */
  SPHDEM_PAIRS_SPH2DEM = 2,  /* calcNeighboursCoupling, SPH side (dataLL.cpp:1154-1190) */
  SPHDEM_PAIRS_DEM2SPH = 3   /* calcNeighboursCoupling, DEM side (dataLL.cpp:1106-1126) */
};

typedef struct sphdem_ctx sphdem_ctx;

/* ---- context ----------------------------------------------------------------
 * Replaces CdataLL::CdataLL(particleContainer&, CglobalVars&, bool)
 * (dataLL.cpp:8-40): binds the constants and the device.  capacity bounds the
 * particle count (MAX_NUM_PARTICLES_PER_CPU). */
int sphdem_create(const sphdem_params* params, int device, int64_t capacity,
                  sphdem_ctx** out);
int sphdem_destroy(sphdem_ctx* ctx);
const char* sphdem_last_error(const sphdem_ctx* ctx);
/* Changes the run-time calibration constants (drag_a, drag_b, dem_k,
 * dem_gamma, viscosity, ...) of an existing context. */
int sphdem_set_params(sphdem_ctx* ctx, const sphdem_params* params);

/* ---- state --------------------------------------------------------------------
 * Uploads n particles (the particleContainer handed to CdataLL).  Every field not
 * given is initialised as the Cparticle constructor does (particle.h:70-135):
 * zero, porosity 1, omega 0.  iam uses the SPHDEM_* type values. */
int sphdem_upload(sphdem_ctx* ctx, int64_t n, const double* r, const double* v,
                  const double* vhat, const double* mass, const double* h,
                  const double* dens, const int32_t* iam);
int sphdem_num_particles(const sphdem_ctx* ctx, int64_t* n);
int sphdem_set_field(sphdem_ctx* ctx, int field, const void* host);
/* sphdem_set_field from device memory of this context's device (n doubles, or
 * n*3 for a vector field, in original order; the caller's writes to it must be
 * complete), for callers whose state lives on the GPU (the adjoint's pair
 * searches); no counterpart in the reference, whose particles live on the
 * host.  Returns once the field is written. */
int sphdem_set_field_device(sphdem_ctx* ctx, int field, const void* dev);
int sphdem_get_field(sphdem_ctx* ctx, int field, void* host);
int sphdem_set_globals(sphdem_ctx* ctx, const sphdem_globals* g);
int sphdem_get_globals(sphdem_ctx* ctx, sphdem_globals* g);
/* Blocks until the context's stream is idle. */
int sphdem_sync(sphdem_ctx* ctx);

/* ---- the three half-steps (sphIncompress.cpp:3, 40, 167) ---------------------- */
int sphdem_step_start(sphdem_ctx* ctx);
int sphdem_step_middle(sphdem_ctx* ctx);
int sphdem_step_end(sphdem_ctx* ctx);

/* ---- device hooks: common bodies of CcustomSim traversals --------------------
 * type_mask bit t selects particles with iam == t (bit 11 = dem). */
/* p.f += accel for selected particles (addGravity, sediment3D/customSim.cpp:5-7) */
int sphdem_hook_body_force(sphdem_ctx* ctx, uint32_t type_mask, const double accel[3]);
/* DEM floor contact at z = RMIN[2] (addCustomBoundaries, sediment3D/customSim.cpp:67-81) */
int sphdem_hook_dem_floor(sphdem_ctx* ctx);
/* freezeParticles (sediment3D/customSim.cpp:12-25) applied when time <= t_freeze */
int sphdem_hook_freeze(sphdem_ctx* ctx, uint32_t type_mask, double t_freeze);
/* sumDragForces (sediment3D/customSim.cpp:27-42): custom[0..5] = sums of m*fdrag */
int sphdem_hook_sum_drag(sphdem_ctx* ctx);

/* Hook sets for sphdem_run_steps. */
#define SPHDEM_HOOKS_NONE 0
#define SPHDEM_HOOKS_DAMBREAK 1  /* gravity on sph before middle()                  */
#define SPHDEM_HOOKS_SEDIMENT 2  /* the sediment3D CcustomSim (customSim.cpp:85-113) */
/* Runs nsteps of run.cpp:102-118: sphStep++, hooks, start/middle/end.  On a
 * step without a density re-initialisation start()'s per-particle work runs
 * inside middle()'s reset (its drift in the "wrap" kernel, its kicks in the
 * "permute" kernel) and end()'s drift inside end()'s reset, with the same
 * arithmetic as the phase calls: the timing window then counts no "start"
 * launches. */
int sphdem_run_steps(sphdem_ctx* ctx, int nsteps, int hookset);

/* Ensemble of independent boxes: advances each of the nbox contexts by nsteps
 * (as sphdem_run_steps on each) with `workers` host threads (0: min(nbox, 8)),
 * each thread stepping its share of the boxes round robin, every box on its
 * own stream, so the small boxes' kernels overlap on the device (during the
 * call each box runs all its passes on its main stream: the boxes already
 * overlap one another).  A box that has run a step gets its steps launched
 * with the resets in their device form (the ghost count stays on the device,
 * and a device check stops the step whose reset needs the host: the cell grid
 * would change, particles were deleted, more ghosts than the launched
 * capacity, an error bit), one step of each of a thread's boxes in turn, and
 * nothing read back until the thread has launched all of them; a stopped box
 * is restored to the step it stopped at (from a device copy of the state kept
 * each step: one more particle store per context) and finishes as
 * sphdem_run_steps.  The reference runs such ensembles as separate processes, one case
 * directory per run (the run.cpp loop once per box); this replaces that loop
 * over boxes.  Every box ends exactly as sphdem_run_steps alone would leave it.
 * The contexts must be distinct and must not be used by the caller during the
 * call.  On an error the first failing box's index goes to *failed_box (if
 * given; -1 otherwise) and its sphdem_last_error holds the message; boxes that
 * had not failed may have advanced by fewer than nsteps steps. */
int sphdem_run_steps_many(sphdem_ctx* const* ctxs, int nbox, int nsteps, int hookset,
                          int workers, int* failed_box);

/* The same steps replayed from a captured HIP graph.  No counterpart in the
 * reference (a CPU loop has no launch cost); results are those of
 * sphdem_run_steps bit for bit.  One step (hooks, start, middle, end) is
 * captured after a step run as launches, with resets that read nothing back to
 * the host: the ghost count stays on the device, bounded by the captured
 * ghost entries, and a device check stops a replay whose reset needs the host
 * (the cell grid would change, particles were deleted, more ghosts than
 * captured, an error bit).  Each replay first keeps a device copy of the state
 * it starts from; a stopped replay's step is restored and run as launches, and
 * the step is captured again.  Steps with a density reinitialisation run as
 * launches.  Memory: one more copy of the particle store per context.  The
 * graph is replayed on the context's stream with no host round trip between
 * replays; the call waits once at its end. */
int sphdem_run_steps_graph(sphdem_ctx* ctx, int nsteps, int hookset);
/* sphdem_run_steps_many with each box's steps replayed from its graph: the
 * worker threads capture where needed, then queue all replays of their boxes
 * on the boxes' streams, then wait for each box once.  Same arguments and
 * results as sphdem_run_steps_many. */
int sphdem_run_steps_many_graph(sphdem_ctx* const* ctxs, int nbox, int nsteps, int hookset,
                                int workers, int* failed_box);
/* sphdem_run_steps_many_graph with one graph over the whole ensemble: every
 * box's captured step a child node of one graph, launched once per ensemble
 * step on one stream (when every box replays the same number of steps and all
 * are on one device; otherwise as sphdem_run_steps_many_graph).  The graph is
 * kept for the next call with the same captures.  Same arguments and results
 * as sphdem_run_steps_many.  No counterpart in the reference. */
int sphdem_run_steps_ensemble_graph(sphdem_ctx* const* ctxs, int nbox, int nsteps, int hookset,
                                    int workers, int* failed_box);

/* ---- parity and per-pass entry points -------------------------------------- */
/* CdataLL::reset() (dataLL.cpp:140-236): periodic wrap, out-of-range deletion,
 * hmax, cell grid hysteresis, binning.  Invalidates cached neighbour sets. */
int sphdem_reset(sphdem_ctx* ctx);
/* getCellI / dem_getCellI (dataLL.h:303-311) of every particle, n*3 int32,
 * SPH cells for non-DEM particles, DEM cells for DEM particles. */
int sphdem_get_cell_coords(sphdem_ctx* ctx, int32_t* ijk);
/* The pair set of one neighbour search as (a, b, image) triples of original
 * indices; image = (sx+1) + 3(sy+1) + 9(sz+1) names the periodic image of b
 * (13 = b itself).  With a == NULL only *count is returned. */
int sphdem_get_pairs(sphdem_ctx* ctx, int kind, int64_t* count, int32_t* a,
                     int32_t* b, int32_t* image, int64_t capacity);
/* sphdem_get_pairs into device memory of this context's device (a, b, image
 * device pointers of capacity entries each), for callers that keep the pair
 * set on the GPU (the adjoint's differentiable step); no counterpart in the
 * reference, whose neighbour lists never leave the host.  Returns once the
 * pairs are written. */
int sphdem_get_pairs_device(sphdem_ctx* ctx, int kind, int64_t* count, int32_t* a,
                            int32_t* b, int32_t* image, int64_t capacity);
/* One operator pass over the current state (neighbour sets of the last reset). */
int sphdem_pass_omega(sphdem_ctx* ctx);      /* neighboursGroup<calcOmega>        */
int sphdem_pass_eos(sphdem_ctx* ctx);        /* traverse<calcPressSpsoundPdr2>     */
int sphdem_pass_force(sphdem_ctx* ctx);      /* neighbours<calcForce> (+ newDt)    */
int sphdem_pass_dddt(sphdem_ctx* ctx);       /* neighbours<calcDddtDudt>           */
int sphdem_pass_density(sphdem_ctx* ctx);    /* neighboursGroup<calcDensity>       */
int sphdem_pass_porosity(sphdem_ctx* ctx);   /* porosity, shepSum, DEM->SPH drag,
                                                 finalisePorosity (sphIncompress.cpp:64-78) */
int sphdem_pass_fpi(sphdem_ctx* ctx);        /* neighboursGroupCoupling<calcFPI>   */
int sphdem_pass_dem_contacts(sphdem_ctx* ctx); /* neighboursGroup<calcDEMContacts> */
/* Global minimum of the accumulated dt candidates (setGlobalTimestep's min). */
int sphdem_reduce_dt(sphdem_ctx* ctx, double* newdt);

/* ---- diagnostics of an output step: CsphIncompress::calcOutputVars
 * (sphIncompress.h:233-319) -- momentum and energies of the sph particles,
 * maxima and density statistics over sph + sphBoundary, averages over the n
 * particles of the container; aveDensFromMass compares each density with a
 * fresh summation over the cached neighbour lists (calcAveDensFromMass).
 * rmsFF, maxFF, eFF, edFFdt, eViscB and eBForce are zero under the supported
 * flag sets (calcEnergies resets the last two). */
typedef struct sphdem_output {
  double linMom[3], angMom, eKE, eViscF, edViscFdt, eViscB, eBForce, eFF, edFFdt;
  double eElast, edElastdt, eElastExact, eTotal;
  double maxV, maxF, maxFF, rmsFF, aveDens, varDens, aveDensFromMass;
  int64_t n, nSph;
} sphdem_output;
int sphdem_output_vars(sphdem_ctx* ctx, sphdem_output* out);

/* ---- full-state checkpoint (no counterpart: the reference's VTK restart
 * keeps positions, velocities, densities and h but not the force sums, dddt,
 * v0/vhat0, porosity, the Shepard sums, the DEM drag of the previous step or
 * dt, io_data_vtk.cpp:20-115, so its resume is not exact) ---------------------
 * Everything a continued run depends on: every per-particle field and type in
 * original-index order, the globals (time, dt, newDt, sphStep, custom[]) and
 * the container state (deletions, search radius).  Restoring into a context
 * created with the same parameters continues the run bit for bit.  The buffer
 * is plain bytes (sphdem_checkpoint_size of them) that the caller may write to
 * a file. */
int sphdem_checkpoint_size(const sphdem_ctx* ctx, int64_t* bytes);
int sphdem_checkpoint_save(sphdem_ctx* ctx, void* host, int64_t capacity);
int sphdem_checkpoint_restore(sphdem_ctx* ctx, const void* host, int64_t bytes);

/* ---- tile shape ----------------------------------------------------------- */
/* No counterpart in the reference (a CPU code has no launch shape).  With
 * seg_particles = 0 the list build runs one lane per target particle (the
 * per-target search); with 8..64 it packs whole cells of one (x,y) column into
 * segments of at most seg_particles, one 64-lane wavefront per segment.  The
 * SPH pair walkers run walker_block (64, 128 or 256) threads per workgroup.
 * Results do not depend on the shape (pair sets and list order are the
 * reference's for every shape); the defaults are 0 and 256. */
int sphdem_set_tile_shape(sphdem_ctx* ctx, int seg_particles, int walker_block);
int sphdem_get_tile_shape(const sphdem_ctx* ctx, int* seg_particles, int* walker_block);
/* Autotuner: times `steps` steps (after one untimed step) of each shape of
 * {0, 64, 48, 32} x {256, 128, 64} from the context's current state, restored
 * from a full-state checkpoint before each trial and after the last, keeps the
 * fastest and returns it (and its ms per step if ms_per_step is not NULL).
 * The simulation state is left as it was. */
int sphdem_tune_tile_shape(sphdem_ctx* ctx, int steps, int hookset, int* seg_particles,
                           int* walker_block, double* ms_per_step);

/* Number of particles deleted as out of range since the context was created. */
int sphdem_num_deleted(const sphdem_ctx* ctx, int64_t* n);

/* ---- measurement (no counterpart in the reference; its timing is the
 * wtTotal* wall-clock fields of CglobalVars, globalVars.h:163-188) ------------
 * While a timing window is open, every launch of the named kernel ("force",
 * "omega", "dddt", "coupling", "shepsum", "fpi", "contacts", "eos", "start",
 * "kick", "rest", "bin", "rank", "permute", "wrap", "lists" = the SPH
 * neighbour-list build) is bracketed by HIP events
 * on the context's stream; sphdem_timing_end synchronises and returns the
 * summed kernel time and the number of launches timed.  Under
 * VAR_H_CORRECTION the steps sum omega inside their "force" and "dddt"
 * launches, so "omega" times only sphdem_pass_omega. */
int sphdem_timing_begin(sphdem_ctx* ctx, const char* kernel);
int sphdem_timing_end(sphdem_ctx* ctx, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* SPHDEM_C_H */
