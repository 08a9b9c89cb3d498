/*
 * emc_oracle.h -- CPU restatement of ViennaEMC's per-particle bulk EMC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Every function restates the reference algorithm in plain C (FP64, same
 * operation order) and cites the reference file:line it follows.  Paths are
 * relative to the ViennaEMC source tree.
 *
 * Units follow the reference: k in 1/m, energy in eV, times in s, SI force.
 */
#ifndef EMC_ORACLE_H
#define EMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants: include/emcConstants.hpp:9-25 (non-CODATA values) ---- */
#define ORC_PI 3.14159265358979323846
#define ORC_Q 1.60219e-19
#define ORC_KB 1.38066e-23
#define ORC_HBAR 1.05459e-34
#define ORC_EPS0 8.85419e-12
#define ORC_ME 9.11e-31

#define ORC_MAX_DEG 8
#define ORC_MAX_Z 8
#define ORC_MAX_VALLEYS 4
#define ORC_MAX_MECH 32
#define ORC_MAX_BATHS 4

/* ---- random numbers ---------------------------------------------------- */
/* std::mt19937_64 (include/emcUtil.hpp:15), the engine is fixed by the C++
 * standard [rand.predef]. */
typedef struct {
  uint64_t mt[312];
  int mti;
} orc_mt64;

void orc_mt64_seed(orc_mt64 *g, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64 *g);
/* fills out[n] with the next n raw outputs */
void orc_mt64_fill(uint64_t seed, uint64_t *out, size_t n);
/* fills out[p*stride + i], i < n, with the raw stream of the per-particle
 * engine seeded by orc_particle_seed(seed, p0 + p) */
void orc_mt64_fill_particles(uint64_t seed, uint64_t p0, size_t np,
                             uint64_t *out, size_t n, size_t stride);
uint64_t orc_splitmix64(uint64_t x);
/* per-particle seed used by the replay fixtures: splitmix64(seed ^ pid) */
uint64_t orc_particle_seed(uint64_t seed, uint64_t pid);

/* libstdc++ generate_canonical<double,53> on a 64-bit engine:
 * double(x) * 2^-64, clamped below 1 (bits/random.tcc:3348-3378). */
double orc_canonical(uint64_t x);

/* RNG source: a live engine or a recorded buffer of raw draws */
typedef struct {
  int kind; /* 0 = mt19937_64, 1 = recorded buffer */
  orc_mt64 mt;
  const uint64_t *buf;
  size_t pos, len;
  uint64_t ndraws;
  int overflow;
} orc_rng;

void orc_rng_init_mt(orc_rng *r, uint64_t seed);
void orc_rng_init_buf(orc_rng *r, const uint64_t *buf, size_t len);
uint64_t orc_rng_raw(orc_rng *r);
/* std::uniform_real_distribution<double>(a, b)(rng) */
double orc_rng_uniform(orc_rng *r, double a, double b);

/* Philox-4x32-10 (Salmon et al., SC'11), used by the production mode. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

/* ---- valleys: include/ValleyTypes/ ------------------------------------ */
enum {
  ORC_PARA_ISO = 0,      /* emcParabolicIsotropValley.hpp */
  ORC_NONPARA_ISO = 1,   /* emcNonParabolicIsotropValley.hpp */
  ORC_PARA_ANISO = 2,    /* emcParabolicAnisotropValley.hpp */
  ORC_NONPARA_ANISO = 3  /* emcNonParabolicAnistropValley.hpp */
};

typedef struct {
  int kind;
  double relMass[3]; /* isotropic kinds use relMass[0] */
  double particleMass;
  int deg;
  double alpha; /* ignored for parabolic kinds */
  double bottom;
  double rot[ORC_MAX_DEG][9];
  /* derived */
  double effMassCond, effMassDOS, vogt[3];
} orc_valley;

void orc_valley_init(orc_valley *v, int kind, const double relMass[3],
                     double particleMass, int deg, double alpha,
                     double bottom);
/* setSubValleyEllipseCoordSystem (rows normalised) */
void orc_valley_set_axes(orc_valley *v, int sub, const double d1[3],
                         const double d2[3], const double d3[3]);
double orc_valley_energy(const orc_valley *v, const double k[3]);
double orc_valley_gamma(const orc_valley *v, double E);
double orc_valley_normk(const orc_valley *v, double E);
double orc_valley_mcond(const orc_valley *v, double E);
double orc_valley_mdos(const orc_valley *v, double E);
double orc_valley_np(const orc_valley *v);
void orc_valley_to_ellipse(const orc_valley *v, int sub, const double in[3],
                           double out[3]);
void orc_valley_to_device(const orc_valley *v, int sub, const double in[3],
                          double out[3]);
void orc_valley_velocity(const orc_valley *v, const double k[3], double E,
                         int sub, double vel[3]);

/* ---- particle: include/emcParticle.hpp:10-18 --------------------------- */
typedef struct {
  double k[3];
  double energy;
  double tau;
  double grainTau;
  int valley;
  int subValley;
  int region;
} orc_particle;

/* initRandomDirection (emcUtil.hpp:131-139) */
void orc_init_random_direction(double norm, double rand1, double rand2,
                               double res[3]);

/* emcParticleDrift.hpp:11-35 */
void orc_drift(double dt, orc_particle *p, const orc_valley *v, double *pos,
               int dim, const double force[3]);

/* ---- scatter mechanisms: include/ScatterMechanisms/ ------------------- */
enum {
  ORC_MECH_ACOUSTIC = 0,
  ORC_MECH_ZERO_IV_AB = 1,
  ORC_MECH_ZERO_IV_EM = 2,
  ORC_MECH_FIRST_IV_AB = 3,
  ORC_MECH_FIRST_IV_EM = 4,
  ORC_MECH_FROEHLICH_AB = 5,
  ORC_MECH_FROEHLICH_EM = 6,
  /* emcHotPhononFroehlichMechanism.hpp:39-118 / :122-200 */
  ORC_MECH_HOT_FROEHLICH_AB = 7,
  ORC_MECH_HOT_FROEHLICH_EM = 8,
  /* emcScreenedFroehlich{Absorption,Emission}3D
   * (emcScreenedFroehlichInteraction.hpp:96-204): the equilibrium Froehlich
   * rate and angle with the vertex 1/q^2 screened to 1/(q^2 + q_s^2), q_s^2
   * the model's (a shared emcPlasmonScreening, emcPlasmonScreening.hpp) */
  ORC_MECH_SCR_FROEHLICH_AB = 9,
  ORC_MECH_SCR_FROEHLICH_EM = 10,
  ORC_MECH_SCR_HOT_FROEHLICH_AB = 11,
  ORC_MECH_SCR_HOT_FROEHLICH_EM = 12
};

typedef struct {
  int kind;
  int valley;
  int finalValley;
  double scatterConst;
  double phononEnergy;
  double deltaValley;
  int Z;
  int finalSub[ORC_MAX_DEG][ORC_MAX_Z];
  double effMass; /* Froehlich: relEffMass * me */
  double nBose;   /* Froehlich */
  int bath;       /* hot-phonon Froehlich: index of the model's phonon bath */
  /* screened hot-phonon Froehlich: q-resolved occupation in the rate and
   * q-resolved angle (emcScreenedFroehlichInteraction.hpp:225-231) */
  int qResolved, qResolvedAngle;
} orc_mech;

/* ---- LO phonon bath: include/emcPhononBath.hpp ------------------------
 * Non-equilibrium occupation N_q per wavevector bin with event tallies
 * (:56-105), the optional acoustic (Klemens) and TO (Ridley) reservoirs and
 * an optional per-bin lifetime profile.  The tallies are written through
 * the pointers, so a const model can still record events. */
typedef struct {
  int nBins;
  double dq, tauLO, N0, Vsim;
  double *Nq, *nEm, *nAbs;
  double *tauProfile; /* NULL = uniform tauLO */
  double loE, latT;
  int hasAc;
  double acE, tauAc, Nac, NacEq;
  int hasRidley;
  double wR, toE, tauTO, Nto, NtoEq;
  /* prefix sums of the transition weight q/(q^2 + qs2Screen) and of
   * weight * N_q, nBins + 1 entries each (:73-82, 109-124) */
  double qs2Screen;
  double *cumW, *cumWN;
  /* energy record of the last update [eV in the bath volume]: loE x the
   * net emissions, loE x sum_q Dph (N_new - N_old), loE x sum_q Dph
   * (dt / tau_q) (N_old - N*) (the LO energy passed to the reservoirs) */
  double energy[3];
} orc_bath;

/* constructor (:160-192) */
void orc_bath_init(orc_bath *b, int nBins, double dq, double tauLO,
                   double phononEnergy, double latticeTemp, double Vsim,
                   int enableAcoustic, double acPhononEnergy,
                   double tauAcoustic, double wRidley, double toPhononEnergy,
                   double tauTO);
void orc_bath_free(orc_bath *b);
/* binOf (:222-225) */
int orc_bath_bin(const orc_bath *b, double q);
/* recordEmission / recordAbsorption (:237-253) */
void orc_bath_record(const orc_bath *b, double q, int emission);
/* update (:264-358) */
void orc_bath_update(orc_bath *b, double dt);
/* getMeanNq (:461-470) */
double orc_bath_mean_nq(const orc_bath *b);
/* rebuildWindowSums (:109-124), setScreeningQ2 (:203-208) */
void orc_bath_rebuild_window_sums(orc_bath *b);
void orc_bath_set_screening_q2(orc_bath *b, double qs2);
/* getNqInWindow (:378-394), sampleQ (:423-458) */
double orc_bath_nq_in_window(const orc_bath *b, double qMin, double qMax);
double orc_bath_sample_q(const orc_bath *b, double qMin, double qMax, int emission, double r);
/* sampleQResolvedCosTheta (emcScreenedFroehlichInteraction.hpp:87-97) */
double orc_qres_cos_theta(const orc_bath *b, double kI, double kF, int emission, double r);
/* getAcousticTemp / getTOTemp (:480-494) */
double orc_bath_acoustic_temp(const orc_bath *b);
double orc_bath_to_temp(const orc_bath *b);

void orc_mech_acoustic(orc_mech *m, int valley, double sigma, double rho,
                       double vSound, double T);
/* kind = ORC_MECH_{ZERO,FIRST}_IV_{AB,EM}; finalSub[deg][Z] row-major */
void orc_mech_iv(orc_mech *m, int kind, int valley, int finalValley,
                 double defPot, double phononEnergy, double rho, double T,
                 int Z, int initDeg, const int *finalSub);
void orc_mech_froehlich(orc_mech *m, int kind, int valley,
                        double phononEnergy, double relEffMass, double epsHi,
                        double epsLo, double T);
/* kind = ORC_MECH_HOT_FROEHLICH_{AB,EM} bound to bath index `bath`
 * (emcHotPhononFroehlichMechanism.hpp:61-70) */
void orc_mech_hot_froehlich(orc_mech *m, int kind, int valley,
                            double phononEnergy, double relEffMass,
                            double epsHi, double epsLo, int bath);

/* ---- model: particle type + scatter handler (single region) ----------- */
typedef struct {
  int nValleys;
  orc_valley valley[ORC_MAX_VALLEYS];
  int nMech;
  orc_mech mech[ORC_MAX_MECH];
  int nLevels;
  double maxEnergy, dE;
  /* per valley: registered mechanisms in order and normalised cumulative
   * tables [nTables][nLevels] */
  int nTables[ORC_MAX_VALLEYS];
  int tableMech[ORC_MAX_VALLEYS][ORC_MAX_MECH];
  double *cum[ORC_MAX_VALLEYS];
  double tau0[ORC_MAX_VALLEYS];
  int nBaths;
  orc_bath bath[ORC_MAX_BATHS];
  /* what the runners do at the end of every step when baths exist:
   * bathUpdate = 1: orc_model_bath_step, rebuilding the tables on steps
   * whose 1-based count is a multiple of rebuildEvery (the reference's
   * reinitEvery, hotPhononGa2O3.cpp:104-109,280); 0: tallies only */
  int bathUpdate, rebuildEvery;
  long long stepCount;
  /* emcElectron's initEnergyEV (emcElectron.hpp:27,84-85): > 0 selects the
   * fixed-energy initialisation in orc_run_shared */
  double initEnergy;
  /* q_s^2 [1/m^2] of the screened mechanisms (emcPlasmonScreening::getQs2) */
  double qs2;
  /* plasma screening: q_s^2 refreshed from the carriers after every step
   * (updateScreening of examples/hotCarrierMHP/hotCarrierMHP.cpp:563-581)
   * with static permittivity scrEps in the volume scrVol */
  int scrPlasma;
  double scrEps, scrVol;
  /* periodic box [m] of tracked positions: the device's maxPos of the bulk
   * handler (basicBulkParticleHandler.hpp:600-613) */
  double box[3];
} orc_model;

void orc_model_init(orc_model *m, int nLevels, double maxEnergy);
void orc_model_free(orc_model *m);
int orc_model_add_valley(orc_model *m, const orc_valley *v);
/* also mirrors the mutable deltaValley update of the IV mechanisms */
int orc_model_add_mech(orc_model *m, const orc_mech *mech);
double orc_mech_rate(const orc_model *m, orc_mech *mech, double E);
/* fillScatterTables + renormalizeTables (emcScatterHandler.hpp:220-288);
 * raw rates optionally returned in rawOut[valley][table][level] */
void orc_model_build_tables(orc_model *m);
int orc_energy_level(const orc_model *m, double E);
/* returns table index selected (>=0), -1 self-scattering, -2 no interval
 * matched, -3 no tables */
int orc_scatter(const orc_model *m, orc_particle *p, orc_rng *rng);
void orc_mech_scatter(const orc_model *m, const orc_mech *mech,
                      orc_particle *p, orc_rng *rng);
double orc_new_tau(const orc_model *m, int valley, orc_rng *rng);
int orc_model_has_hot(const orc_model *m);
/* end of a time step of the reference's hot-phonon drivers
 * (examples/hotPhononGa2O3/hotPhononGa2O3.cpp:274-281,
 *  examples/hotCarrierMHP/hotCarrierMHP.cpp:693-703): every bath's
 * update(dt), then reinitScatterTables (emcScatterHandler.hpp:100-108)
 * when `rebuild` is set and a mechanism reads a bath */
void orc_model_bath_step(orc_model *m, double dt, int rebuild);
/* the runners' end-of-step hook described at orc_model.bathUpdate; with
 * plasma screening (scrPlasma) it refreshes q_s^2 from the carriers after
 * the bath update and rebuilds the tables every step: nPart carriers, of
 * which cnt0 in valley 0 with energies summing to sumE0 (particle order) */
void orc_model_end_step(orc_model *m, double dt, long long nPart, double sumE0,
                        long long cnt0);
/* the hotCarrierMHP updateScreening (:563-581) for one carrier species:
 * qs2 = n q^2 / (eps eps0 kB T), n = nPart / scrVol, T = getMBTemp of the
 * valley-0 mean energy (emcHotCarrierOutput.hpp:89-92); 0 without carriers
 * or temperature.  Sets the model's qs2 and every bath's setScreeningQ2. */
void orc_model_update_screening(orc_model *m, long long nPart, double sumE0, long long cnt0);
/* valley-0 energy sum and count in particle order (getAvgEnergy,
 * basicBulkParticleHandler.hpp:304-321) */
void orc_valley0_sums(int np, const double *E, const int32_t *valley, double *sumE0,
                      long long *cnt0);
void orc_valley0_sums_aos(int np, const orc_particle *p, double *sumE0, long long *cnt0);

/* ---- bulk loop: examples/bulkSimulation/basicBulkParticleHandler.hpp --- */
typedef struct {
  int32_t nEvents;     /* scatter calls (incl. self scattering) */
  int32_t nReal;       /* events that selected a mechanism */
  int32_t lastMech;    /* table index of the last real event, -1 if none */
} orc_step_info;

/* one particle, one step (:195-221); evCount[nTables] optional */
void orc_move_particle(const orc_model *m, orc_particle *p, double tStep,
                       const double force[3], orc_rng *rng,
                       orc_step_info *info, int32_t *evCount, double *pos);

/* Replay runner: np particles with per-particle raw streams
 * draws[p*stride ...].  Arrays are SoA, k[3*np] (x,y,z per particle).
 * Outputs per step (optional, may be NULL): valley/sub/nEvents [nsteps*np],
 * kOut [nsteps*np*3], EOut [nsteps*np].  Returns number of particles that
 * ran out of draws. */
int orc_run_replay(orc_model *m, int np, double *k, double *E,
                   double *tau, double *grainTau, int32_t *valley,
                   int32_t *sub, const uint64_t *draws, size_t stride,
                   size_t *cursor, int nsteps, double tStep,
                   const double force[3], int32_t *valleyOut,
                   int32_t *subOut, int32_t *nEventsOut, double *kOut,
                   double *EOut, int32_t *evCount, double *pos, double *posOut);

/* per-particle initial state from a per-particle stream
 * (emcElectron.hpp:75-90, emcParticleInitialization.hpp:36-51), with the
 * bulk coordinate taken as an interior node (no boundary flips). */
void orc_init_particle(const orc_model *m, orc_particle *p, double Vt,
                       orc_rng *rng);
/* the same with initParticleKSpaceFixed (emcParticleInitialization.hpp:
 * 57-71): E = E0, isotropic k (emcElectron.hpp:84-85) */
void orc_init_particle_fixed(const orc_model *m, orc_particle *p, double E0,
                             orc_rng *rng);

/* ensemble observables (basicBulkParticleHandler.hpp:289-347), per valley */
void orc_observables(const orc_model *m, int np, const orc_particle *parts,
                     const double fieldDir[3], double *avgE, double *avgV,
                     double *occ);

/* Full restatement of the reference bulk driver with one shared
 * mt19937_64 stream (single thread): generateInitialParticles (:143-159)
 * on a 3-D box, then nsteps of moveParticles with the per-step
 * observables of examples/bulkSimulation/bulkSimulation.cpp:150-156.
 * obs[(step)*(3*nValleys)] = avgE[v], avgV[v], occ[v] for step 0..nsteps;
 * with baths, each bath's getMeanNq after every step goes to
 * bathMeanOut[(step-1)*nBaths + b] (may be NULL).
 * Returns the number of particles created (or -needed if > maxParticles).*/
int orc_run_shared(orc_model *m, uint64_t seed, const double maxPos[3],
                   const double spacing[3], double doping, double T,
                   const double fieldDir[3], double fieldStrength,
                   double tStep, int nsteps, int maxParticles, double *obs,
                   double *kOut, double *EOut, int32_t *valleyOut,
                   int32_t *subOut, double *bathMeanOut, double *posInitOut,
                   double *posOut);

/* ---- production mode (Philox draws, FP64) ------------------------------
 * The CPU twin of the GPU production kernel: same algorithm as
 * orc_move_particle, with random numbers drawn from Philox-4x32-10 at
 * counter (event, step, particle, instance) and the final subvalley drawn
 * from the conditional uniform of the selection draw.  Units: SI. */
void orc_move_particle_philox(const orc_model *m, orc_particle *p,
                              double tStep, const double force[3],
                              uint64_t seed, uint32_t instance, uint32_t pid,
                              uint32_t step, orc_step_info *info,
                              int32_t *evCount, double *pos);
void orc_init_particle_philox(const orc_model *m, orc_particle *p, double Vt,
                              uint64_t seed, uint32_t instance, uint32_t pid);
/* one production-mode scatter event e of the (step, particle) stream:
 * selection + final state, returns the next free-flight time [s] */
double orc_philox_event(const orc_model *m, orc_particle *p, uint64_t seed,
                        uint32_t inst, uint32_t pid, uint32_t step, uint32_t e);
/* uniform j of the (step, particle) stream of an instance word */
double orc_prod_uniform(uint64_t seed, uint32_t inst, uint32_t pid, uint32_t step,
                        uint32_t j);
/* production uniform of one Philox word: (w >> 8 | 1) * 2^-24, strictly
 * inside (0, 1), the GPU kernels' value in both precisions */
double orc_philox_u01_24(uint32_t w);


/* ---- handle API for language bindings (tests) -------------------------- */
orc_model *orc_model_new(int nLevels, double maxEnergy);
void orc_model_delete(orc_model *m);
/* axes: deg*9 doubles (rows dir1, dir2, dir3 per subvalley) or NULL */
int orc_model_add_valley_params(orc_model *m, int kind, double m0, double m1,
                                double m2, double particleMass, int deg,
                                double alpha, double bottom,
                                const double *axes);
int orc_model_add_acoustic(orc_model *m, int valley, double sigma, double rho,
                           double vSound, double T);
int orc_model_add_iv(orc_model *m, int kind, int valley, int finalValley,
                     double defPot, double phononEnergy, double rho, double T,
                     int Z, const int *finalSub);
int orc_model_add_froehlich(orc_model *m, int kind, int valley,
                            double phononEnergy, double relEffMass,
                            double epsHi, double epsLo, double T);
/* kind = ORC_MECH_SCR_FROEHLICH_{AB,EM} (constructor :107-122 / :163-178) */
int orc_model_add_screened_froehlich(orc_model *m, int kind, int valley,
                                     double phononEnergy, double relEffMass,
                                     double epsHi, double epsLo, double T);
/* emcPlasmonScreening::setQs2 (enabled): the q_s^2 the screened rates and
 * angles read (rebuild the tables afterwards) */
void orc_model_set_qs2(orc_model *m, double qs2);
/* kind = ORC_MECH_SCR_HOT_FROEHLICH_{AB,EM} (constructors :238-250 /
 * :313-325) bound to bath index `bath` */
int orc_model_add_screened_hot_froehlich(orc_model *m, int kind, int valley,
                                         double phononEnergy, double relEffMass,
                                         double epsHi, double epsLo, int bath,
                                         int qResolved, int qResolvedAngle);
/* plasma screening on (eps, volume) or off (eps <= 0) */
void orc_model_set_screening_plasma(orc_model *m, double eps, double volume);
/* refresh q_s^2 from the SoA state and rebuild the tables: the seeding
 * before the first step (hotCarrierMHP.cpp:630-637) */
void orc_model_seed_screening(orc_model *m, int np, const double *E, const int32_t *valley);
void orc_model_get_bath_window(const orc_model *m, int b, double *cumW, double *cumWN);
double orc_model_get_qs2(const orc_model *m);
double orc_bath_nq_window_one(const orc_model *m, int b, double qMin, double qMax);
double orc_bath_sample_q_one(const orc_model *m, int b, double qMin, double qMax, int emission,
                             double r);
double orc_qres_cos_theta_one(const orc_model *m, int b, double kI, double kF, int emission,
                              double r);
/* screenedFroehlichLogFactor / sampleScreenedFroehlichCosTheta
 * (emcScreenedFroehlichInteraction.hpp:52-77) */
double orc_scr_log_factor(double kI, double kF, double qs2);
double orc_scr_cos_theta(double kI, double kF, double qs2, double r);
int orc_model_n_tables(const orc_model *m, int valley);
int orc_model_table_mech(const orc_model *m, int valley, int j);
double orc_model_tau0(const orc_model *m, int valley);
void orc_model_get_cum(const orc_model *m, int valley, double *out);
double orc_model_rate(orc_model *m, int mech, double E);
double orc_model_dE(const orc_model *m);
/* one scatter event (selection + final state) on a particle given as
 * arrays; returns orc_scatter's code; draws consumed in *ndraws */
int orc_scatter_one(const orc_model *m, double *k, double *E, int32_t *valley,
                    int32_t *sub, const uint64_t *draws, size_t ndraw,
                    size_t *ndrawsUsed);
/* one mechanism's final state from a raw stream */
void orc_mech_scatter_one(const orc_model *m, int mech, double *k, double *E,
                          int32_t *valley, int32_t *sub,
                          const uint64_t *draws, size_t ndraw,
                          size_t *ndrawsUsed);
void orc_drift_one(const orc_model *m, int valley, int sub, double dt,
                   double *k, double *E, double *pos, const double *force);
void orc_velocity_one(const orc_model *m, int valley, int sub,
                      const double *k, double E, double *vel);
double orc_normk_one(const orc_model *m, int valley, double E);
/* per-particle init from per-particle raw streams (draws[p*stride..]) */
void orc_init_particles_replay(const orc_model *m, int np, double Vt,
                               const uint64_t *draws, size_t stride,
                               size_t *cursor, double *k, double *E,
                               double *tau, double *grainTau, int32_t *valley,
                               int32_t *sub);
/* production-mode twin over np particles (SoA, SI units); evCount may be
 * NULL; obs[(s)*3*nv ...] = avgE, avgV, occ per step (may be NULL) */
void orc_run_philox(orc_model *m, int np, double *k, double *E,
                    double *tau, int32_t *valley, int32_t *sub, int nsteps,
                    double tStep, const double force[3],
                    const double fieldDir[3], uint64_t seed,
                    uint32_t instance, uint32_t step0, double *obs,
                    int32_t *nEventsOut, double *pos);
void orc_init_particles_philox(const orc_model *m, int np, double Vt,
                               uint64_t seed, uint32_t instance, double *k,
                               double *E, double *tau, int32_t *valley,
                               int32_t *sub, double *pos);
/* the periodic box [m] of tracked positions */
void orc_model_set_box(orc_model *m, const double box[3]);
void orc_init_particle_philox_energy(const orc_model *m, orc_particle *p,
                                     double E0, double sigma, uint64_t seed,
                                     uint32_t inst, uint32_t pid);
void orc_init_particles_philox_energy(const orc_model *m, int np, double E0,
                                      double sigma, uint64_t seed,
                                      uint32_t instance, double *k, double *E,
                                      double *tau, int32_t *valley,
                                      int32_t *sub);
/* phonon baths and hot-phonon mechanisms */
int orc_model_add_bath(orc_model *m, int nBins, double dq, double tauLO,
                       double phononEnergy, double latticeTemp, double Vsim,
                       int enableAcoustic, double acPhononEnergy,
                       double tauAcoustic, double wRidley,
                       double toPhononEnergy, double tauTO);
void orc_model_set_bath_profile(orc_model *m, int b, const double *profile);
int orc_model_add_hot_froehlich(orc_model *m, int kind, int valley,
                                double phononEnergy, double relEffMass,
                                double epsHi, double epsLo, int bath);
void orc_model_get_bath(const orc_model *m, int b, double *Nq, double *nEm,
                        double *nAbs, double *scal);
void orc_model_set_bath(orc_model *m, int b, const double *Nq, double Nac,
                        double Nto);
void orc_model_bath_record_one(orc_model *m, int b, double q, int emission);
void orc_model_bath_update_one(orc_model *m, int b, double dt);
/* the energy record of bath b's last update (orc_bath.energy) */
void orc_model_get_bath_energy(const orc_model *m, int b, double out[3]);
void orc_model_set_bath_mode(orc_model *m, int update, int rebuildEvery,
                             long long stepCount);
void orc_model_set_init_energy(orc_model *m, double E0);
/* mt19937_64 ensemble (reference RNG, one stream per particle) with
 * per-step observables; the statistical baseline for production mode.
 * nThreads > 1 runs particles in parallel (pthreads); the wall time of
 * the step loop alone (no initialisation) goes to *stepSeconds; subOcc
 * (may be NULL): per step the fraction in each (valley, subvalley),
 * [step][valley][ORC_MAX_DEG]. */
void orc_run_mt(orc_model *m, int np, uint64_t seed, double Vt,
                int nsteps, double tStep, const double force[3],
                const double fieldDir[3], double *obs, int nThreads,
                double *kOut, double *EOut, int32_t *valleyOut,
                int32_t *subOut, double *stepSeconds, double *subOcc);

#ifdef __cplusplus
}
#endif
#endif /* EMC_ORACLE_H */
