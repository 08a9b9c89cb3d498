/*
 * emcgpu.h -- C ABI of the MI355X Ensemble Monte Carlo particle step.
 *
 * Plain C, status-code returns, no exceptions across the boundary, opaque
 * handle.  One context per host thread / HIP stream; a context is not
 * internally thread-safe.  All SI quantities at the boundary use the
 * reference's units (k in 1/m, energy in eV, time in s, field in V/m).
 *
 * Reference interfaces replaced (paths relative to the ViennaEMC tree):
 *   - particle type + scatter handler:
 *       include/ParticleType/emcParticleType.hpp:97-200,
 *       include/emcScatterHandler.hpp:79-289
 *   - valley plugins:   include/ValleyTypes/emcAbstractValley.hpp:20-91
 *   - mechanism plugins: include/ScatterMechanisms/emcScatterMechanism.hpp:17-53
 *   - bulk handler: examples/bulkSimulation/basicBulkParticleHandler.hpp:93-347
 *   - free flight:  include/emcParticleDrift.hpp:11-35
 */
#ifndef EMCGPU_H
#define EMCGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMCGPU_MAX_VALLEYS 4
#define EMCGPU_MAX_DEG 8
#define EMCGPU_MAX_Z 8
#define EMCGPU_MAX_MECH 32
#define EMCGPU_MAX_LEVELS 8192
#define EMCGPU_MAX_BATHS 4
#define EMCGPU_MAX_BINS 1024

typedef struct emcgpu_ctx emcgpu_ctx;

typedef enum {
  EMCGPU_OK = 0,
  EMCGPU_ERR_ARG = 1,     /* invalid argument (message says which) */
  EMCGPU_ERR_HIP = 2,     /* HIP runtime error */
  EMCGPU_ERR_STATE = 3,   /* call out of order (e.g. no tables yet) */
  EMCGPU_ERR_DRAWS = 4,   /* replay: a particle ran out of recorded draws */
  EMCGPU_ERR_NOMEM = 5
} emcgpu_status;

typedef enum { EMCGPU_FP64 = 0, EMCGPU_FP32 = 1 } emcgpu_precision;
typedef enum { EMCGPU_MODE_PRODUCTION = 0, EMCGPU_MODE_REPLAY = 1 } emcgpu_mode;

/* valley kinds: include/ValleyTypes/emc{Parabolic,NonParabolic}{Isotrop,
 * Anis(o)trop}Valley.hpp */
typedef enum {
  EMCGPU_VALLEY_PARA_ISO = 0,
  EMCGPU_VALLEY_NONPARA_ISO = 1,
  EMCGPU_VALLEY_PARA_ANISO = 2,
  EMCGPU_VALLEY_NONPARA_ANISO = 3
} emcgpu_valley_kind;

/* mechanism kinds (rate formula + final-state sampler):
 *   ACOUSTIC      emcAcousticScatterMechanism.hpp:47-72
 *   ZERO_IV_*     emcZeroOrderInterValleyScatterMechanism.hpp:11-21,103-129,246-272
 *   FIRST_IV_*    emcFirstOrderInterValleyScatterMechanism.hpp:11-20,102-131,250-279
 *   FROEHLICH_*   emcFroehlichInteraction.hpp:45-57,95-128,167-201
 *   HOT_FROEHLICH_* emcHotPhononFroehlichMechanism.hpp:39-200: the Froehlich
 *                 rate with the occupation of a phonon bath (getMeanNq) and
 *                 the Froehlich final state, recording |k' - k| in the bath
 *   SCREENED_FROEHLICH_* emcScreenedFroehlichInteraction.hpp:52-77,101-213:
 *                 the Froehlich rate with the vertex 1/q^2 screened to
 *                 1/(q^2 + q_s^2) and the angle drawn from the screened
 *                 distribution; q_s^2 from emcgpu_set_screening
 *   SCREENED_HOT_FROEHLICH_* emcScreenedFroehlichInteraction.hpp:87-97,
 *                 218-375: the screened rate with the occupation of a phonon
 *                 bath, either its mean (getMeanNq) or, with q_resolved, its
 *                 coupling-weighted mean over the reachable window
 *                 [|kI - kF|, kI + kF] (getNqInWindow, emcPhononBath.hpp:
 *                 378-394); the angle screened, or with q_resolved and
 *                 q_resolved_angle drawn through the bath's q distribution
 *                 (sampleQ, :423-458); |k' - k| recorded in the bath
 *   custom        rate supplied by the caller (emcgpu_set_rate_tables) with
 *                 one of the built-in final states */
typedef enum {
  EMCGPU_MECH_ACOUSTIC = 0,
  EMCGPU_MECH_ZERO_IV_AB = 1,
  EMCGPU_MECH_ZERO_IV_EM = 2,
  EMCGPU_MECH_FIRST_IV_AB = 3,
  EMCGPU_MECH_FIRST_IV_EM = 4,
  EMCGPU_MECH_FROEHLICH_AB = 5,
  EMCGPU_MECH_FROEHLICH_EM = 6,
  EMCGPU_MECH_HOT_FROEHLICH_AB = 7,
  EMCGPU_MECH_HOT_FROEHLICH_EM = 8,
  EMCGPU_MECH_SCREENED_FROEHLICH_AB = 9,
  EMCGPU_MECH_SCREENED_FROEHLICH_EM = 10,
  EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB = 11,
  EMCGPU_MECH_SCREENED_HOT_FROEHLICH_EM = 12
} emcgpu_mech_kind;

typedef struct {
  int device;      /* HIP device ordinal; -1 = host-only context (model,
                      tables and validation, no device memory) */
  int precision;   /* emcgpu_precision of the particle state and kinematics */
  int n_instances; /* independent (bias, temperature, material) instances */
} emcgpu_config;

/* emcAbstractValley subclasses, lowered to a descriptor */
typedef struct {
  int kind;              /* emcgpu_valley_kind */
  double rel_mass[3];    /* relative masses; isotropic kinds use [0] */
  double particle_mass;  /* kg (emcParticleType::getMass) */
  int degeneracy;        /* number of subvalleys */
  double alpha;          /* non-parabolicity [1/eV]; ignored if parabolic */
  double bottom_energy;  /* eV */
  double axes[EMCGPU_MAX_DEG][9]; /* per subvalley: dir1, dir2, dir3 rows of
                                     setSubValleyEllipseCoordSystem; need not
                                     be normalised; anisotropic kinds only */
} emcgpu_valley_desc;

/* emcScatterMechanism subclasses, lowered to a descriptor */
typedef struct {
  int kind;           /* emcgpu_mech_kind */
  int valley;         /* initial valley (getIdxValley) */
  int final_valley;   /* IV kinds */
  double phonon_energy; /* eV (IV, Froehlich) */
  double coupling;    /* acoustic: deformation potential [eV]; zero-order IV:
                         [eV/m]; first-order IV: [eV]; unused otherwise */
  double rel_eff_mass; /* Froehlich kinds: m* / m0 of the rate prefactor */
  double eps_hi, eps_lo; /* Froehlich: relative permittivities */
  int z;              /* IV: number of final subvalleys per initial one */
  int final_sub[EMCGPU_MAX_DEG][EMCGPU_MAX_Z]; /* IV: finalSubValleys map */
  int bath;           /* HOT_FROEHLICH_*, SCREENED_HOT_FROEHLICH_*: index of the
                         instance's phonon bath */
  int q_resolved;     /* SCREENED_HOT_FROEHLICH_*: window occupation in the rate */
  int q_resolved_angle; /* SCREENED_HOT_FROEHLICH_* with q_resolved: angle from
                           the bath's q distribution (the reference's default,
                           inQResolvedAngle = true) */
} emcgpu_mech_desc;

/* emcPhononBath constructor + setTauLOProfile (include/emcPhononBath.hpp:
 * 160-192, 214-216): LO phonon occupation N_q on n_bins wavevector bins,
 * optionally coupled to an acoustic (Klemens) and a TO (Ridley) reservoir */
typedef struct {
  int n_bins;              /* <= EMCGPU_MAX_BINS */
  double dq;               /* bin width [1/m] */
  double tau_lo;           /* LO lifetime [s] */
  double phonon_energy;    /* hbar w_LO [eV] */
  double lattice_temp;     /* [K] */
  double volume;           /* simulated volume V_sim [m^3] */
  int acoustic;            /* non-zero: enable the acoustic reservoir */
  double ac_phonon_energy; /* hbar w_ac [eV] */
  double tau_acoustic;     /* [s] */
  double w_ridley;         /* Ridley branching ratio (0 = off) */
  double to_phonon_energy; /* hbar w_TO [eV] */
  double tau_to;           /* [s] */
  const double *tau_profile; /* per-bin LO lifetime [s] (n_bins), or NULL */
} emcgpu_bath_desc;

typedef struct {
  double eps_r, rho, ni, v_sound; /* emcMaterial.hpp:23-25 (rho kg/m^3, v m/s) */
} emcgpu_material;

/* emcPlasmonScreening (include/emcPlasmonScreening.hpp:52-90): the q_s^2
 * that an instance's SCREENED_FROEHLICH_* mechanisms share.
 *   FIXED   q_s^2 as given (setQs2).  Like the reference's setQs2 it acts on
 *           the final-state angle at once; the rates follow at the next
 *           emcgpu_build_tables (reinitScatterTables).
 *   PLASMA  refreshed on the GPU after every step from the instance's
 *           carriers, q_s^2 = n q^2 / (eps_static eps0 kB T) with
 *           n = N / volume and T = 2 <E>_valley0 q / (3 kB) (getMBTemp), 0
 *           without carriers; then every bath's window weights and the
 *           screened tables are rebuilt (the updateScreening and
 *           reinitScatterTables of examples/hotCarrierMHP/hotCarrierMHP.cpp:
 *           563-581, 693-703).  emcgpu_update_screening seeds it before the
 *           first step (:630-637).
 * Every bath of the instance weights its window sums with the same q_s^2
 * (setScreeningQ2). */
typedef enum { EMCGPU_SCREENING_FIXED = 0, EMCGPU_SCREENING_PLASMA = 1 } emcgpu_screening_mode;

typedef struct {
  int mode;          /* emcgpu_screening_mode */
  int enabled;       /* 0: q_s^2 stays 0 (emcPlasmonScreening(eps, false)),
                        the screened kinds reduce to the bare Froehlich ones */
  double eps_static; /* static relative permittivity (Debye form) */
  double qs2;        /* q_s^2 [1/m^2], >= 0 (PLASMA: until the first update) */
  double volume;     /* PLASMA: volume of the carriers [m^3] */
} emcgpu_screening_desc;

typedef struct {
  double dt;            /* time step [s] */
  int n_steps;          /* steps to advance */
  int steps_per_launch; /* steps fused per kernel launch (0 = the value of
                           emcgpu_autotune, else min(n_steps, 64)) */
  int mode;             /* emcgpu_mode */
  uint64_t seed;        /* production: Philox key */
  uint32_t step0;       /* production: global index of the first step */
  int timing;           /* non-zero: time every launch with HIP events */
  int particles_per_lane; /* production kernel: particles owned by each
                             GPU lane, 1, 2, 4 or 8 (0 = the value of
                             emcgpu_autotune, else the library default) */
  int use_graph;        /* production: capture the call's launches (step and
                           reduction kernels) in a HIP graph and launch it;
                           the executable graph is updated in place on the
                           next call.  Ignored with timing or trace. */
  /* instances with phonon baths: after every step each bath is updated
   * (emcPhononBath::update) and the tables of instances whose mechanisms
   * read a bath are rebuilt from it on the GPU, as the reference's
   * hot-phonon drivers do (examples/hotPhononGa2O3/hotPhononGa2O3.cpp:
   * 274-281).  Steps are then launched one at a time. */
  int bath_record_only;  /* non-zero: only tally the events (no update) */
  int rebuild_every;     /* rebuild after steps whose 1-based global index
                            step0 + s + 1 is a multiple of this (reinitEvery
                            of hotPhononGa2O3.cpp:104-109); 0 or 1 = every */
  /* replay (FP64 contexts only): per-particle recorded raw 64-bit draws of
   * the reference engine, host memory, particle order of instance 0 */
  const uint64_t *draws; /* [n_particles][draw_stride] */
  size_t draw_stride;
  uint64_t *cursor;      /* [n_particles] in/out: next draw to consume */
  double *grain_tau;     /* [n_particles] in/out seconds, may be NULL */
  /* optional per-step trace of instance 0 (host, may be NULL) */
  int32_t *trace_valley; /* [n_steps][n] */
  int32_t *trace_sub;    /* [n_steps][n] */
  int32_t *trace_events; /* [n_steps][n] scatter calls incl. self scattering */
  double *trace_k;       /* [n_steps][n][3] 1/m */
  double *trace_energy;  /* [n_steps][n] eV */
  double *trace_pos;     /* [n_steps][n][3] m, with tracked positions (may be NULL) */
} emcgpu_step_args;

/* ---- lifecycle ---------------------------------------------------------
 * Replaces constructing emcParticleType + basicBulkParticleHandler
 * (basicBulkParticleHandler.hpp:93-120). */
emcgpu_status emcgpu_create(const emcgpu_config *cfg, emcgpu_ctx **out);
void emcgpu_destroy(emcgpu_ctx *ctx);
/* last error message of ctx (or of the failing create when ctx is NULL) */
const char *emcgpu_last_error(const emcgpu_ctx *ctx);

/* ---- model --------------------------------------------------------------
 * addValley / addScatterMechanism (emcParticleType.hpp:125-153): valleys in
 * index order, mechanisms in registration order (table order per valley). */
emcgpu_status emcgpu_set_model(emcgpu_ctx *ctx, int instance,
                               const emcgpu_valley_desc *valleys, int n_valleys,
                               const emcgpu_mech_desc *mechs, int n_mechs);

/* initScatterTables (emcScatterHandler.hpp:91-95, 220-288) with the built-in
 * rate formulas, FP64 on the host; region 0. */
emcgpu_status emcgpu_build_tables(emcgpu_ctx *ctx, int instance, int n_levels,
                                  double max_energy, double temperature,
                                  const emcgpu_material *material);

/* Same, with caller-supplied raw rates rates[mech][level] at energies
 * (level+1)*max_energy/n_levels, e.g. from a custom getScatterRate. */
emcgpu_status emcgpu_set_rate_tables(emcgpu_ctx *ctx, int instance,
                                     int n_levels, double max_energy,
                                     const double *rates);

/* initScatterTables of every instance at once (emcScatterHandler.hpp:220-288)
 * with the rates of the built-in kinds evaluated on the GPU from the
 * descriptors, one lane per (level, mechanism); temperature[n_instances],
 * mat[n_instances].  The FP64 formulas are the host's, the device's pow /
 * sqrt / log within an ulp of them.  Instances whose rates follow a bath or
 * the carriers use emcgpu_build_tables. */
emcgpu_status emcgpu_build_tables_all(emcgpu_ctx *ctx, int n_levels, double max_energy,
                                      const double *temperature, const emcgpu_material *mat);

/* normalised cumulative tables cum[table][level] and tau0 [s] of a valley
 * (emcParticleType::getTau, emcParticleType.hpp:181-183) */
emcgpu_status emcgpu_get_tables(const emcgpu_ctx *ctx, int instance,
                                int valley, double *cum, double *tau0,
                                int *n_tables);

/* applied field: force = normalize(dir) * strength * charge
 * (basicBulkParticleHandler.hpp:100-102, resetAppliedFieldStrength :135) */
emcgpu_status emcgpu_set_field(emcgpu_ctx *ctx, int instance,
                               const double dir[3], double strength,
                               double charge);

/* ---- particles ----------------------------------------------------------
 * emcParticle SoA storage (include/emcParticle.hpp:10-18) */
emcgpu_status emcgpu_resize(emcgpu_ctx *ctx, const size_t *counts);
emcgpu_status emcgpu_get_count(const emcgpu_ctx *ctx, int instance,
                               size_t *count);
emcgpu_status emcgpu_set_particles(emcgpu_ctx *ctx, int instance,
                                   const double *k, const double *energy,
                                   const double *tau, const int32_t *valley,
                                   const int32_t *sub);
emcgpu_status emcgpu_get_particles(const emcgpu_ctx *ctx, int instance,
                                   double *k, double *energy, double *tau,
                                   int32_t *valley, int32_t *sub);
/* thermal initial state on the GPU (emcElectron::generateInitialParticle,
 * emcElectron.hpp:75-90 + emcParticleInitialization.hpp:36-51), Philox */
emcgpu_status emcgpu_init_particles(emcgpu_ctx *ctx, int instance,
                                    uint64_t seed, double temperature);
/* initial state at a prescribed energy: sigma = 0 is emcElectron's
 * fixed-energy photoexcitation (initEnergyEV > 0: emcElectron.hpp:84-85,
 * initParticleKSpaceFixed, emcParticleInitialization.hpp:57-71); sigma > 0
 * draws E from N(mean, sigma) by Box-Muller on two Philox uniforms
 * (|E - mean| <= 5.26 sigma, |E| taken for a negative draw). */
emcgpu_status emcgpu_init_particles_energy(emcgpu_ctx *ctx, int instance,
                                           uint64_t seed, double mean_ev,
                                           double sigma_ev);

/* ---- positions of bulk particles ------------------------------------------
 * The bulk handler keeps a position per particle, drifted with the free
 * flight (include/emcParticleDrift.hpp:26-34) and wrapped into the device
 * box [0, maxPos] (basicBulkParticleHandler.hpp:600-613); it is written by
 * print (:227-248).  emcgpu_set_box turns this on for the context (the
 * step kernels then run one particle per lane) with the instance's box
 * max_pos[3] [m]; emcgpu_init_particles then also draws positions uniform
 * in the box.  Positions are [n][3] in m. */
emcgpu_status emcgpu_set_box(emcgpu_ctx *ctx, int instance, const double max_pos[3]);
emcgpu_status emcgpu_set_positions(emcgpu_ctx *ctx, int instance, const double *pos);
emcgpu_status emcgpu_get_positions(const emcgpu_ctx *ctx, int instance, double *pos);

/* ---- plasmon screening ---------------------------------------------------
 * Replaces the shared_ptr<emcPlasmonScreening> of the screened mechanisms
 * (emcScreenedFroehlichInteraction.hpp:105, 114) and its setQs2/getQs2
 * (emcPlasmonScreening.hpp:79-82).  Per instance; kept across
 * emcgpu_set_model; q_s^2 = 0 until set. */
emcgpu_status emcgpu_set_screening(emcgpu_ctx *ctx, int instance,
                                   const emcgpu_screening_desc *desc);
emcgpu_status emcgpu_get_screening(const emcgpu_ctx *ctx, int instance, double *qs2);
/* PLASMA: q_s^2 from the instance's current particles (sum of valley-0
 * energies in particle order, like getAvgEnergy) and the screened tables
 * rebuilt: the seeding before the first step */
emcgpu_status emcgpu_update_screening(emcgpu_ctx *ctx, int instance);

/* ---- phonon baths --------------------------------------------------------
 * Baths of an instance (replacing the shared_ptr<emcPhononBath> given to
 * the hot-phonon mechanisms, emcHotPhononFroehlichMechanism.hpp:45); set
 * them before emcgpu_build_tables.  State starts at N_q = N0. */
emcgpu_status emcgpu_set_baths(emcgpu_ctx *ctx, int instance,
                               const emcgpu_bath_desc *baths, int n_baths);
/* current state: nq[n_bins] (may be NULL) and scalars[6] = getMeanNq,
 * getMeanNac, getMeanNTO, getN0, getAcousticTemp, getTOTemp
 * (emcPhononBath.hpp:461-501) (may be NULL) */
emcgpu_status emcgpu_get_bath(const emcgpu_ctx *ctx, int instance, int bath,
                              double *nq, double *scalars);
/* overwrite N_q and the reservoir occupations (tables follow on the next
 * emcgpu_step) */
emcgpu_status emcgpu_set_bath_state(emcgpu_ctx *ctx, int instance, int bath,
                                    const double *nq, double n_ac,
                                    double n_to);
/* event tallies not yet consumed by an update (recordEmission /
 * recordAbsorption, emcPhononBath.hpp:237-253), exact integers */
emcgpu_status emcgpu_get_bath_tallies(const emcgpu_ctx *ctx, int instance,
                                      int bath, uint64_t *n_em,
                                      uint64_t *n_abs);
/* getMeanNq after every step of the last emcgpu_step: mean_nq[n_steps] */
emcgpu_status emcgpu_get_bath_history(const emcgpu_ctx *ctx, int instance,
                                      int bath, double *mean_nq,
                                      int *n_steps);
/* energy record of every bath update of the last emcgpu_step,
 * energy[step][3] in eV within the bath volume, FP64 sums in bin order over
 * the update of emcPhononBath.hpp:264-358: [0] hbar w_LO x (emissions -
 * absorptions) of the step, [1] hbar w_LO x sum_q D_ph(q) (N_q after - N_q
 * before) (LO energy gained), [2] hbar w_LO x sum_q D_ph(q) dt / tau(q)
 * (N_q before - N*) (LO energy passed to the reservoirs); [0] = [1] + [2]
 * unless the clamp N_q >= 0 acted.  No counterpart in the reference. */
emcgpu_status emcgpu_get_bath_energy(const emcgpu_ctx *ctx, int instance,
                                     int bath, double *energy, int *n_steps);

/* ---- time stepping ------------------------------------------------------
 * moveParticles(dt) repeated n_steps times
 * (basicBulkParticleHandler.hpp:181-225) with the per-step observables of
 * examples/bulkSimulation/bulkSimulation.cpp:150-156 reduced on the GPU. */
emcgpu_status emcgpu_step(emcgpu_ctx *ctx, const emcgpu_step_args *args);

/* Sizes the step buffers (per-step partials, observables, reduction pairs,
 * bath history) and uploads the model for a later emcgpu_step with the
 * same arguments (or fewer steps), which then allocates nothing; launches
 * nothing.  No counterpart in the reference (std::vector::reserve). */
emcgpu_status emcgpu_reserve(emcgpu_ctx *ctx, const emcgpu_step_args *args);

/* emcgpu_step without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * per-step partials and reduction pairs use the library's static scratch
 * (64 MiB per GPU, held until the process ends); the per-step observables
 * (layout of emcgpu_get_observables, all instances: obs[instance][step][j])
 * go to the caller's device buffer obs of obs_len doubles, which must hold
 * n_instances * n_steps * n_obs values.  Production mode only: no replay,
 * trace or use_graph, no phonon baths or plasma screening, and the model
 * must already be on the device for args->dt (an earlier emcgpu_step or
 * emcgpu_init_particles with that dt after the last model, table or field
 * change).  Steps per launch are capped so a launch's partials fit the
 * scratch; an ensemble whose single step does not fit is refused
 * (EMCGPU_ERR_ARG).  Results equal emcgpu_step's.  The scratch is shared by
 * every context on a GPU and keeps the last call's contents: calls that
 * overlap on one GPU (other threads or streams) race. */
emcgpu_status emcgpu_step_noalloc(emcgpu_ctx *ctx, const emcgpu_step_args *args,
                                  double *obs, size_t obs_len);

/* Per-step ensemble sums of the last emcgpu_step, FP64:
 * out[step][j], j < n_obs, with
 *   j = 0: scatter calls (incl. self scattering); j = 1: real events;
 *   for valley v (< max valleys): 2+3v: count, 3+3v: sum E [eV],
 *                                 4+3v: sum v.dir [m/s];
 *   then per (valley, subvalley): count.
 * Averages of getAvgEnergy / getAvgDriftVelocity /
 * getValleyOccupationProbability (basicBulkParticleHandler.hpp:289-347)
 * are sums / counts. */
emcgpu_status emcgpu_get_observables(const emcgpu_ctx *ctx, int instance,
                                     double *out, int *n_obs, int *n_steps);
emcgpu_status emcgpu_obs_layout(const emcgpu_ctx *ctx, int *n_obs,
                                int *max_valleys, int *max_deg);
/* the ensemble sums of the current state of every instance without a step,
 * out[instance][j] in the layout above (scatter counts 0): the getAvg*
 * values right after generateInitialParticles
 * (bulkSimulation.cpp:143-146) */
emcgpu_status emcgpu_observe(emcgpu_ctx *ctx, double *out);

/* ---- 2-D device simulation ------------------------------------------------
 * The device path of emcSimulation (emcSimulation.hpp:106-193) on instance 0
 * of a context, electrons only (charge -q, one carrier per particle):
 *   geometry        emcDevice (emcDevice.hpp:83-343), emcDopingProfile,
 *                   emcSurface ohmic and gate contacts (emcSurface.hpp:199-408)
 *   charge / force  PMSchemes/emcNGPScheme.hpp:24-66, emcCICScheme.hpp:32-173,
 *                   emcNECScheme.hpp:31-113 (the reference's weights, quirks
 *                   included)
 *   field           PMSchemes/emcEFieldCalculation.hpp:13-81
 *   Poisson         the discrete equations of PoissonSolver/emcSORSolver.hpp:
 *                   49-197, solved by Newton's method with an FP32 multigrid
 *                   inner solve and FP64 residuals (iterative refinement)
 *   particles       emcBasicParticleHandler (ParticleHandler/
 *                   emcBasicParticleHandler.hpp:76-254): drift with positions,
 *                   boundaries, ohmic-contact equilibrium, injection
 * Positions are in m at the boundary; grids are x-fastest (iy * nx + ix),
 * potentials normalised by V_T as in the reference. */
#define EMCGPU_MAX_CONTACTS 8
#define EMCGPU_MAX_REGIONS 16

typedef enum { EMCGPU_PM_NGP = 0, EMCGPU_PM_CIC = 1, EMCGPU_PM_NEC = 2 } emcgpu_pm_scheme;
typedef enum {
  EMCGPU_XMIN = 0, EMCGPU_XMAX = 1, EMCGPU_YMIN = 2, EMCGPU_YMAX = 3
} emcgpu_boundary;

/* addConstantDopingRegion (emcDevice.hpp:162-167), applied in order */
typedef struct {
  double min_pos[2], max_pos[2]; /* [m] */
  double doping;                 /* [1/m^3] */
} emcgpu_doping_region;

/* emcContactType (emcContact.hpp:8); Schottky contacts need the degenerate
 * 2-D carrier model of the reference and are rejected */
typedef enum {
  EMCGPU_CONTACT_OHMIC = 0, EMCGPU_CONTACT_SCHOTTKY = 1, EMCGPU_CONTACT_GATE = 2
} emcgpu_contact_type;

/* addOhmicContact (emcDevice.hpp:177-182) and addGateContact (:195-202).
 * An ohmic contact is a Dirichlet node set marked on every boundary its nodes
 * lie on; a gate adds the Robin term of checkForGateContact
 * (PoissonSolver/emcSORSolver.hpp:399-412) on its own boundary, reflects
 * particles like an artificial boundary and keeps the contact rule of the
 * E-field boundary.  Zero-initialised fields give an ohmic contact. */
typedef struct {
  int boundary;     /* emcgpu_boundary */
  double voltage;   /* [V] */
  double min_pos, max_pos; /* extent along the boundary [m] */
  int type;         /* emcgpu_contact_type */
  double eps_oxide;      /* gate: relative permittivity of the oxide */
  double thickness;      /* gate: oxide thickness [m] */
  double barrier_height; /* gate: barrier height [V] */
} emcgpu_contact;

typedef struct {
  double max_pos[2];   /* device extent [m] */
  double spacing[2];   /* grid spacing [m] */
  double width;        /* device width [m] (setDeviceWidth) */
  double temperature;  /* [K] */
  double eps_r;        /* relative permittivity */
  double ni;           /* intrinsic carrier density [1/m^3] */
  int n_doping;
  const emcgpu_doping_region *doping;
  int n_contacts;
  const emcgpu_contact *contacts;
  int pm_scheme;        /* emcgpu_pm_scheme */
  double poisson_tol;   /* relative residual: max |F| / max h^2 (|C| + n + p) */
  int poisson_max_iter; /* Newton / refinement iterations per solve */
  /* emcElectron's usePotentialForInit (emcElectron.hpp:48-61): initial
   * carriers per node from exp(equilibrium potential) * n_i instead of the
   * doping */
  int init_from_potential;
  /* per-region scatter tables (the (valley, region) table sets of
   * emcScatterHandler, emcScatterHandler.hpp:53-68, 130-145): nonzero makes
   * instance r hold the tables of doping region r (the mechanisms registered
   * for r, addScatterMechanism(regions, ...)); a particle then scatters with
   * the tables, and draws its free-flight times with the tau0, of the region
   * of its position (updated after every drift, emcAbstractParticleHandler.
   * hpp:244-247).  Every node must lie in a doping region, the context needs
   * an instance per region, each with the valleys of instance 0 and its
   * tables built.  0: instance 0's tables everywhere. */
  int region_tables;
} emcgpu_device_desc;

/* geometry (host) and device grids; call after emcgpu_set_model and
 * emcgpu_build_tables of instance 0.  Needs a GPU context except for the
 * geometry getters below.  The Poisson solver runs in one workgroup: at
 * most 1024 nodes along x; grids whose multigrid levels exceed the LDS
 * (above about 11k nodes) keep them in device memory instead. */
emcgpu_status emcgpu_device_setup(emcgpu_ctx *ctx, const emcgpu_device_desc *desc);
/* ext[2] node counts; scal[4] = V_T [V], intrinsic Debye length [m], cell
 * volume [m^3], normalised cell volume (emcDevice.hpp:236-343) */
emcgpu_status emcgpu_device_info(const emcgpu_ctx *ctx, int *ext, double *scal);
/* per node: doping [1/m^3], doping region (-1: none), ohmic contact index
 * (-1: none); each may be NULL */
emcgpu_status emcgpu_device_grids(const emcgpu_ctx *ctx, double *doping, int *region,
                                  int *contact);

/* emcgpu_poisson -- calcEquilibriumPotential (equilibrium != 0, holes and
 * electrons Boltzmann) or calcNonEquilibriumPotential (electrons = conc,
 * normalised by n_i) in place on pot[nodes] (normalised).  reset_bc sets the
 * ohmic nodes first (V/V_T + asinh(C/2), no V in equilibrium).  iters and
 * resid (relative residual reached) may be NULL. */
emcgpu_status emcgpu_poisson(emcgpu_ctx *ctx, double *pot, const double *conc,
                             int equilibrium, int reset_bc, int *iters, double *resid);
/* emcgpu_efield -- calcEField of the context's scheme: E [V/m] at nodes
 * (NGP, CIC) or edge midpoints (NEC) from pot (normalised) */
emcgpu_status emcgpu_efield(emcgpu_ctx *ctx, const double *pot, double *ex, double *ey);
/* emcgpu_deposit -- assignToMesh of n positions pos[n][2] [m]: carriers per
 * node, accumulated into grid[nodes] */
emcgpu_status emcgpu_deposit(emcgpu_ctx *ctx, const double *pos, size_t n, double *grid);
/* interpolateForce of the context's scheme: force[n][2] [N] on a particle
 * of the given charge from ex, ey (V/m, as emcgpu_efield lays them out) */
emcgpu_status emcgpu_interpolate(emcgpu_ctx *ctx, const double *ex, const double *ey,
                                 const double *pos, size_t n, double charge, double *force);

/* ---- 3-D particle-mesh operators ----------------------------------------
 * The reference's 3-D schemes on a caller grid (emcDevice<T, 3> with
 * emcNGPScheme / emcCICScheme; its NEC scheme is 2-D only): ext[0] x ext[1]
 * x ext[2] nodes, x fastest (emcGrid.hpp:241-253), positions [n][3] in
 * metres.  FP64 whatever the context's precision; host buffers; need a GPU
 * context, no device setup. */
typedef struct {
  int ext[3];         /* nodes per dimension */
  double spacing[3];  /* [m] */
  int pm_scheme;      /* EMCGPU_PM_NGP or EMCGPU_PM_CIC */
} emcgpu_grid3;
/* assignToMesh, one carrier per position, accumulated into grid[nodes]
 * (emcNGPScheme.hpp:36-48; emcCICScheme.hpp:97-120, the node at the floor
 * corner taking wX wY wZ); nodes outside the grid are skipped */
emcgpu_status emcgpu_deposit3(emcgpu_ctx *ctx, const emcgpu_grid3 *grid3, const double *pos,
                              size_t n, double *grid);
/* interpolateForce: force[n][3] [N] on a particle of the given charge from
 * the nodal field ex, ey, ez [V/m] (emcNGPScheme.hpp:51-66;
 * emcCICScheme.hpp:143-173, including its (1 - wY)(1 - wY) weight of the
 * (x+1, y+1) nodes) */
emcgpu_status emcgpu_interpolate3(emcgpu_ctx *ctx, const emcgpu_grid3 *grid3, const double *ex,
                                  const double *ey, const double *ez, const double *pos, size_t n,
                                  double charge, double *force);
/* calcEFieldAtGridPts with setEFieldBoundaryValues
 * (emcEFieldCalculation.hpp:13-32, 57-81): ex, ey, ez [V/m] from the
 * potential pot (normalised by the thermal voltage vt [V]); at least 3
 * nodes per dimension.  bc[3][nodes]: for component d, a node on a face
 * normal to d holds -1 (artificial boundary: E_d = 0) or a contact index
 * >= 0 (E_d = the inner neighbour's), every other node -2 (the status
 * emcSurface::isArtificialBoundary gives the face points) */
emcgpu_status emcgpu_efield3(emcgpu_ctx *ctx, const emcgpu_grid3 *grid3, double vt,
                             const signed char *bc, const double *pot, double *ex, double *ey,
                             double *ez);
/* emcgpu_contacts -- the ohmic-contact equilibrium of handleOhmicContacts
 * (emcBasicParticleHandler.hpp:158-192) on positions in array order:
 * keep[n] = 0 for particles deleted as excess, inject[nodes] = particles to
 * inject (generateInjectedParticles, emcAbstractParticleHandler.hpp:200-216) */
emcgpu_status emcgpu_contacts(emcgpu_ctx *ctx, const double *pos, size_t n, int *keep,
                              int *inject);

/* ---- device particles ------------------------------------------------------
 * set / get the ensemble in array order: pos[n][2] [m], k[n][3] [1/m], E [eV],
 * tau [s], valley, sub (any output may be NULL; *n in: capacity of the
 * outputs, out: particles held) */
emcgpu_status emcgpu_device_set_particles(emcgpu_ctx *ctx, size_t n, const double *pos,
                                          const double *k, const double *energy,
                                          const double *tau, const int32_t *valley,
                                          const int32_t *sub);
emcgpu_status emcgpu_device_get_particles(const emcgpu_ctx *ctx, size_t *n, double *pos,
                                          double *k, double *energy, double *tau,
                                          int32_t *valley, int32_t *sub);
/* frozen field for emcgpu_device_drift: ex, ey [V/m] as emcgpu_efield lays
 * them out */
emcgpu_status emcgpu_device_set_efield(emcgpu_ctx *ctx, const double *ex, const double *ey);

typedef struct {
  double dt;            /* [s] */
  int n_steps;
  int mode;             /* emcgpu_mode */
  uint64_t seed;        /* production */
  uint32_t step0;       /* production: first step index of the Philox counter */
  /* replay: per-particle raw draws (stride draw_stride) and cursors, the
   * grain timers [s] (in/out, may be NULL) */
  const uint64_t *draws;
  long long draw_stride;
  uint64_t *cursor;
  double *grain_tau;
  /* outputs (may be NULL): exit_step[n] (-1: still inside), exit_contact[n];
   * per step s and particle i: pos[s][i][2], k[s][i][3], E[s][i], events[s][i] */
  int32_t *exit_step, *exit_contact;
  double *tr_pos, *tr_k, *tr_E;
  int32_t *tr_events;
} emcgpu_device_drift_args;

/* driftScatterParticles (emcBasicParticleHandler.hpp:93-139) for n_steps
 * under the frozen field: a particle that leaves through an ohmic contact
 * stops there and is reported; no contact equilibrium, no deposit */
emcgpu_status emcgpu_device_drift(emcgpu_ctx *ctx, const emcgpu_device_drift_args *args);

/* calcEquilibriumCharacteristics + generateInitialParticles + the first
 * charge assignment (emcSimulation.hpp:139-160, emcElectron.hpp:48-104):
 * equilibrium potential, E-field, particles per node (the reference's
 * count with one Philox Bernoulli draw for the remainder), initial states
 * from Philox; *n_particles gets the ensemble size */
emcgpu_status emcgpu_device_init(emcgpu_ctx *ctx, uint64_t seed, size_t *n_particles);

typedef struct {
  double dt;          /* [s] */
  int n_steps;
  uint64_t seed;
  int use_graph;      /* capture the per-step kernel sequence in a HIP graph */
  /* per-step records (may be NULL): [n_steps][n_contacts] particles that
   * left through / were deleted at / were injected at each contact, the
   * ensemble size after the step, the work of the step's Poisson solve
   * (dense Newton corrections applied + multigrid V-cycles) and its relative
   * residual, and |sum(nodes) - N| / N of the charge assignment the step's
   * Poisson solve used (FP64 check sum) */
  int32_t *exits, *deleted, *injected;
  int64_t *n_particles;
  int32_t *poisson_iters;
  double *poisson_resid, *checksum;
  /* setPoissonInterval (emcSimulation.hpp:84-86): Poisson and E-field on
   * the steps whose index (counted from emcgpu_device_init, 0-based) is a
   * multiple of this, the field kept in between; 0 or 1 = every step.
   * Steps without a solve record 0 iterations and residual 0. */
  int poisson_interval;
  /* setChannelCurrentRegion (emcSimulation.hpp:64-70): with
   * channel_length > 0, channel_current[n_steps] gets after each step
   * q * sum v_x / channel_length [A] over the particles with
   * channel_x0 <= x <= channel_x1 [m] (getChannelDriftCurrent,
   * emcBasicParticleHandler.hpp:224-238) */
  double channel_x0, channel_x1, channel_length;
  double *channel_current;
  /* non-zero: add each step's potential and concentration to the running
   * sums read by emcgpu_device_averages (updateAverageCharacteristics,
   * emcSimulationResults.hpp:85-93) */
  int accumulate;
  /* how the per-step Poisson solves reach the stopping rule of
   * emcgpu_poisson (the solution is the same to that tolerance): up to two
   * Newton corrections from an inverse of the discrete operator without the
   * hole term, each after a residual check, then the multigrid-
   * preconditioned Newton solve checks the residual and finishes the solve
   * if needed.  The inverse is
   *   separable: for devices whose Dirichlet nodes are whole columns and
   *     rows and whose gates span whole boundary lines, the operator is a
   *     Kronecker sum of two 1-D operators; their FP64 eigenbases (rocSOLVER,
   *     once per geometry) apply the inverse exactly in O(n (nx + ny)),
   *   dense: otherwise a dense FP32 matrix of the active nodes (rocSOLVER
   *     Cholesky, once per geometry), streamed with FP64 sums over all CUs,
   *     for up to 24576 active nodes.
   * Corrections run from the second step on, while the hole term h^2 p
   * stays below 1e-9 of the diagonal.
   * EMCGPU_POISSON_AUTO: separable, else dense, else multigrid alone;
   * EMCGPU_POISSON_DENSE: dense, else multigrid alone;
   * EMCGPU_POISSON_MULTIGRID: multigrid alone. */
  int poisson_mode;
} emcgpu_device_step_args;

#define EMCGPU_POISSON_AUTO 0
#define EMCGPU_POISSON_MULTIGRID 1
#define EMCGPU_POISSON_DENSE 2

/* performEMCStep repeated n_steps times (emcSimulation.hpp:164-193):
 * Poisson, E-field, driftScatterParticles, handleOhmicContacts, injection,
 * charge assignment; the step counter carries over between calls */
emcgpu_status emcgpu_device_step(emcgpu_ctx *ctx, const emcgpu_device_step_args *args);

/* current potential (normalised), E-field [V/m], electron concentration
 * (normalised) and carriers per node of the last assignment (each may be
 * NULL) */
emcgpu_status emcgpu_device_fields(const emcgpu_ctx *ctx, double *pot, double *ex,
                                   double *ey, double *conc, double *nr_part);
/* averages of the potential and concentration (normalised) over the steps
 * run with accumulate != 0 since emcgpu_device_init or the last reset
 * (avgPot / avgConc of writeFinalResults, emcSimulationResults.hpp:175-183);
 * *n_steps gets the number of steps; reset != 0 clears the sums afterwards */
emcgpu_status emcgpu_device_averages(emcgpu_ctx *ctx, double *pot, double *conc,
                                     long long *n_steps, int reset);

/* accumulated kernel time of timed emcgpu_step calls (HIP events on the
 * context's stream): total ms and number of particle-step launches */
emcgpu_status emcgpu_get_timing(const emcgpu_ctx *ctx, double *ms_total,
                                long long *launches, int reset);
emcgpu_status emcgpu_synchronize(emcgpu_ctx *ctx);

/* ---- launch-shape autotuner (no counterpart in the reference, whose
 * OpenMP loop has no launch shape) ------------------------------------------
 * Times production runs of the context's current ensemble (all instances)
 * for every pair of candidate particles-per-lane (tile = 256 x ppl
 * particles) and steps fused per launch, restoring the particle state
 * before every trial and at the end, and keeps the fastest pair as the
 * context's default for emcgpu_step arguments left at 0.  Results of
 * emcgpu_step do not depend on the launch shape (random numbers are
 * indexed by particle and step), so tuning changes only speed.  The
 * observables of the last emcgpu_step are discarded.  Not available with
 * phonon baths (their state evolves with every step); emcgpu_resize clears
 * the tuned values. */
#define EMCGPU_TUNE_MAX 32
typedef struct {
  double dt;            /* step of the trial runs [s] */
  int n_steps;          /* steps per trial (0 = twice the largest
                           steps-per-launch candidate) */
  uint64_t seed;        /* Philox key of the trial runs */
  uint32_t step0;
  int repeats;          /* timed trials per candidate after one untimed
                           warm-up; the fastest counts (0 = 1) */
  int n_ppl;            /* candidates (0: 1, 2, 4, 8) */
  int ppl[4];
  int n_spl;            /* candidates (0: 16, 32, 64, 128) */
  int spl[8];
  /* results */
  int n_results;
  int res_ppl[EMCGPU_TUNE_MAX], res_spl[EMCGPU_TUNE_MAX];
  double res_ms_per_step[EMCGPU_TUNE_MAX];
  int best_ppl, best_spl;
  double best_ms_per_step;
} emcgpu_autotune_args;
emcgpu_status emcgpu_autotune(emcgpu_ctx *ctx, emcgpu_autotune_args *args);

#ifdef __cplusplus
}
#endif
#endif /* EMCGPU_H */
