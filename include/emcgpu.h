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
  EMCGPU_MECH_HOT_FROEHLICH_EM = 8
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
  double rel_eff_mass; /* Froehlich: m* / m0 of the rate prefactor */
  double eps_hi, eps_lo; /* Froehlich: relative permittivities */
  int z;              /* IV: number of final subvalleys per initial one */
  int final_sub[EMCGPU_MAX_DEG][EMCGPU_MAX_Z]; /* IV: finalSubValleys map */
  int bath;           /* HOT_FROEHLICH_*: index of the instance's phonon bath */
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

typedef struct {
  double dt;            /* time step [s] */
  int n_steps;          /* steps to advance */
  int steps_per_launch; /* steps fused per kernel launch (0 = automatic) */
  int mode;             /* emcgpu_mode */
  uint64_t seed;        /* production: Philox key */
  uint32_t step0;       /* production: global index of the first step */
  int timing;           /* non-zero: time every launch with HIP events */
  int particles_per_lane; /* production kernel: particles owned by each
                             GPU lane, 1, 2, 4 or 8 (0 = library default,
                             or the autotuned value) */
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

/* ---- time stepping ------------------------------------------------------
 * moveParticles(dt) repeated n_steps times
 * (basicBulkParticleHandler.hpp:181-225) with the per-step observables of
 * examples/bulkSimulation/bulkSimulation.cpp:150-156 reduced on the GPU. */
emcgpu_status emcgpu_step(emcgpu_ctx *ctx, const emcgpu_step_args *args);

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

/* accumulated kernel time of timed emcgpu_step calls (HIP events on the
 * context's stream): total ms and number of particle-step launches */
emcgpu_status emcgpu_get_timing(const emcgpu_ctx *ctx, double *ms_total,
                                long long *launches, int reset);
emcgpu_status emcgpu_synchronize(emcgpu_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* EMCGPU_H */
