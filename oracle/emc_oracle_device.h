/*
 * emc_oracle_device.h -- CPU restatement of ViennaEMC's 2-D device path:
 * grid, doping regions and ohmic contacts, the particle-mesh schemes (NGP,
 * CIC, NEC), the electric field, the SOR Poisson solver, the device particle
 * loop with boundary handling, ohmic-contact equilibrium and the
 * emcSimulation step order.
 *
 * TEST INFRASTRUCTURE ONLY.  Like emc_oracle.h this is the checker for the
 * HIP product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Every function restates the reference in
 * plain C (FP64, same operation order) and cites the file:line it follows
 * (paths relative to the ViennaEMC source tree).
 *
 * Grids are stored x-fastest, index = iy * ext[0] + ix (emcGrid.hpp:242-254).
 */
#ifndef EMC_ORACLE_DEVICE_H
#define EMC_ORACLE_DEVICE_H

#include "emc_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_CONTACTS 8
#define ORC_MAX_REGIONS 16

enum { ORC_XMIN = 0, ORC_XMAX = 1, ORC_YMIN = 2, ORC_YMAX = 3, ORC_BINVALID = 4 };
enum { ORC_PM_NGP = 0, ORC_PM_CIC = 1, ORC_PM_NEC = 2 };

typedef struct {
  int ext[2];
  double maxPos[2], h[2], width, T;
  double epsR, ni;
  double Vt, debye, cellVolume, normCellVolume;
  double *doping; /* SI, initialised with ni (emcDopingProfile.hpp:59-63) */
  int *region;    /* doping region index, -1 where none was added */
  int nRegions;
  double regionDoping[ORC_MAX_REGIONS];
  int *contactGrid[4]; /* per boundary XMIN, XMAX (ext[1] nodes), YMIN, YMAX (ext[0]) */
  int nContacts;
  double contactV[ORC_MAX_CONTACTS];
  /* emcContactType (emcContact.hpp:8): 0 ohmic, 2 gate; gate parameters
   * (emcGateContact, emcContact.hpp:80-100): oxide permittivity, oxide
   * thickness [m], barrier height [V] */
  int contactType[ORC_MAX_CONTACTS];
  double gateEps[ORC_MAX_CONTACTS], gateThick[ORC_MAX_CONTACTS], gateBarrier[ORC_MAX_CONTACTS];
  /* per doping region, the particle type's tables of that region (the
   * (valley, region) table sets of emcScatterHandler, emcScatterHandler.hpp:
   * 53-68, 130-145): when set, scattering and getNewTau of a particle use
   * regionModel[particle.region]; otherwise the run's single model */
  const orc_model *regionModel[ORC_MAX_REGIONS];
  int nRegionModels;
} orc_device;

/* the model of a particle in region r (the run's model m without region
 * models) */
const orc_model *orc_device_model_at(const orc_device *d, const orc_model *m, int region);
/* region models, n <= ORC_MAX_REGIONS (n = 0: none) */
int orc_device_set_region_models(orc_device *d, const orc_model *const *models, int n);

/* emcDevice ctor (emcDevice.hpp:83-99), maxPosToExtent (emcUtil.hpp:112-121) */
void orc_device_init(orc_device *d, double maxX, double maxY, double hx,
                     double hy, double T, double epsR, double ni);
void orc_device_free(orc_device *d);
/* setDeviceWidth + calcCellVolume (emcDevice.hpp:104-109, 333-343) */
void orc_device_set_width(orc_device *d, double width);
/* addConstantDopingRegion (emcDevice.hpp:162-167, emcDopingProfile.hpp:74-81,
 * 161-207) */
void orc_device_add_doping(orc_device *d, const double minPos[2],
                           const double maxPos[2], double doping);
/* addOhmicContact (emcDevice.hpp:177-182, emcSurface.hpp:205-233, 381-408) */
int orc_device_add_ohmic(orc_device *d, int boundary, double voltage,
                         double minPos, double maxPos);
/* addGateContact (emcDevice.hpp:195-202, emcSurface.hpp:198-214): marked on
 * its own boundary only */
int orc_device_add_gate(orc_device *d, int boundary, double voltage, double minPos,
                        double maxPos, double epsOxide, double thickness,
                        double barrierHeight);
/* posToCoord (emcDevice.hpp:290-297) */
void orc_device_pos_to_coord(const orc_device *d, const double pos[2],
                             size_t coord[2]);
int orc_device_is_out_of_bounds(const orc_device *d, const double pos[2]);
/* getBoundaryPos / getContactIdx / isOhmicContact (emcSurface.hpp:92-99,
 * 251-259, 328-337) */
int orc_device_boundary_of(const orc_device *d, const size_t coord[2]);
int orc_device_contact_idx(const orc_device *d, const size_t coord[2]);
int orc_device_is_ohmic(const orc_device *d, const size_t coord[2]);
size_t orc_device_idx(const orc_device *d, size_t ix, size_t iy);

/* ---- particle-mesh schemes (PMSchemes/) ------------------------------- */
/* assignToMesh over many positions (pos[2*n]); grid is accumulated */
void orc_pm_assign(int scheme, const double *pos, int n, double nrCarriers,
                   const double h[2], const int ext[2], double *grid);
/* interpolateForce; eField = [Ex grid, Ey grid] */
void orc_pm_force(int scheme, const double *ex, const double *ey,
                  const double pos[2], const double h[2], const int ext[2],
                  double charge, double force[3]);
/* calcEField: NGP / CIC at grid points, NEC at edge midpoints
 * (PMSchemes/emcEFieldCalculation.hpp:13-81) */
void orc_pm_efield(int scheme, const orc_device *d, const double *pot,
                   double *ex, double *ey);

/* 3-D schemes on a grid of ext[0] x ext[1] x ext[2] nodes (x fastest),
 * spacing h [m]; NGP or CIC (the reference's NEC is 2-D only).
 * orc_pm3_assign: assignToMesh of n positions pos[3n], one carrier each,
 * accumulated into grid.  orc_pm3_force: interpolateForce into force[3n]
 * from E = ex, ey, ez.  orc_pm3_efield: calcEFieldAtGridPts with vt the
 * thermal voltage and bc[3][nodes] the boundary status of each node for
 * the component normal to the faces it lies on (-2 on none, -1 artificial,
 * >= 0 contact) */
void orc_pm3_assign(int scheme, const int ext[3], const double h[3], const double *pos, int n,
                    double *grid);
void orc_pm3_force(int scheme, const int ext[3], const double h[3], const double *ex,
                   const double *ey, const double *ez, const double *pos, int n, double charge,
                   double *force);
void orc_pm3_efield(const int ext[3], const double h[3], double vt, const int *bc,
                    const double *pot, double *ex, double *ey, double *ez);

/* ---- SOR Poisson solver (PoissonSolver/emcSORSolver.hpp) --------------- */
typedef struct {
  double accuracy; /* normalised (accuracy [V] / Vt) */
  double omega;
  double hNorm[2], hFactor[2], hFactorSum, hProduct;
  size_t maxCoord[2];
  double *dopingNorm;
  long long sweeps; /* sweeps of the last solve */
  /* gate characteristics per contact (initDeviceCharacteristics,
   * emcSORSolver.hpp:355-366): 2 (eps_ox / eps_sc) / t_ox, the normalised
   * barrier height and applied voltage */
  double gateFactor[ORC_MAX_CONTACTS], gateBarrier[ORC_MAX_CONTACTS],
      gateVoltage[ORC_MAX_CONTACTS];
} orc_sor;

void orc_sor_init(orc_sor *s, const orc_device *d, double accuracyV,
                  double omega);
void orc_sor_free(orc_sor *s);
/* calcEquilibriumPotential (:49-128, beta = 0) */
void orc_sor_equilibrium(orc_sor *s, const orc_device *d, double *pot,
                         int resetBC);
/* calcNonEquilibriumPotential (:131-197, beta = 0); eConc normalised */
void orc_sor_nonequilibrium(orc_sor *s, const orc_device *d, double *pot,
                            const double *eConc, int resetBC);
/* initPotential (emcSimulationResults.hpp:208-213) */
void orc_device_init_potential(const orc_device *d, double *pot);
/* updateCurrentParticleConcentrations (emcSimulationResults.hpp:98-116) */
void orc_device_concentration(const orc_device *d, const double *nrPart,
                              double *conc);

/* ---- device particle loop ------------------------------------------- */
typedef struct {
  orc_particle *p;
  double *pos; /* [n][2] */
  int n, cap;
} orc_ensemble;

/* one particle, one step of emcBasicParticleHandler::driftScatterParticles
 * (ParticleHandler/emcBasicParticleHandler.hpp:93-139) with driftParticle
 * (emcAbstractParticleHandler.hpp:238-249) and handleParticleAtBoundary
 * (emcParticleDrift.hpp:41-66, specular boundaries emcScatterHandler.hpp:
 * 172-189).  Returns 1 if the particle left through an ohmic contact; the
 * contact index goes to *contact. */
int orc_device_move_particle(const orc_model *m, const orc_device *d, int pm,
                             const double *ex, const double *ey, double charge,
                             orc_particle *p, double pos[2], double tStep,
                             orc_rng *rng, int *contact, int32_t *nEvents);

/* A full device simulation in the reference's order (emcSimulation.hpp:
 * 106-193, emcAbstractParticleHandler.hpp:133-148, 200-216,
 * emcBasicParticleHandler.hpp:76-192) with one shared mt19937_64 stream
 * (serial build of the reference).  Per step s (0-based) the outputs hold:
 * nPartOut[s] particles after the step, potOut[s][nodes] the potential used
 * in the step (normalised), exitOut / delOut / injOut[s][contacts] the
 * particles that left, were deleted as excess, were injected.  eqPot gets
 * the equilibrium potential, nInit the initial particle count. */
typedef struct {
  orc_device *dev;
  orc_model *model;
  int pm;
  double accuracyV, omega;
  double tStep;
  int nSteps;
  uint64_t seed;
  double nrCarriers;
  /* setPoissonInterval (emcSimulation.hpp:84-86): Poisson + E-field on
   * steps s with s % poissonInterval == 0 (<= 1: every step) */
  int poissonInterval;
  /* getChannelDriftCurrent (emcBasicParticleHandler.hpp:224-238) after
   * every step into chanOut[s] when chanOut != NULL: region chan[0] <= x <=
   * chan[1] [m], length chan[2] [m] */
  double chan[3];
  double *chanOut;
  /* usePotentialForInit (emcElectron.hpp:48-61): initial carriers from
   * exp(equilibrium potential) * ni instead of the doping */
  int initFromPotential;
} orc_device_run_cfg;

int orc_device_run(const orc_device_run_cfg *cfg, double *eqPot, int *nInit,
                   int *nPartOut, double *potOut, int *exitOut, int *delOut,
                   int *injOut, double *nrPartLast, double *stateOut,
                   int maxState);

/* ohmic-contact equilibrium on a given ensemble: count particles at
 * reservoir nodes in array order, delete the excess, report per contact
 * node the number kept and the number to inject (emcBasicParticleHandler.hpp:
 * 158-192 with generateInjectedParticles, emcAbstractParticleHandler.hpp:
 * 200-216).  keepOut[i] = 1 if particle i survives; injOut[node] = particles
 * to inject at grid node (0 elsewhere); expOut[node] = expected number. */
void orc_device_contact_rule(const orc_device *d, const double *pos, int n,
                             double nrCarriers, int *keepOut, int *injOut,
                             double *expOut);

/* ---- production twin of the GPU device path (Philox streams) -----------
 * orc_device_move_particle with the scatter events of the (step, pid)
 * stream of the GPU kernels, and orc_device_run with Philox creation,
 * initial-count and move streams (stateOut rows: x, y, k, E, tau, valley,
 * sub, pid).  Vt: thermal voltage of the particle initialisation. */
int orc_device_move_particle_philox(const orc_model *m, const orc_device *d, int pm,
                                    const double *ex, const double *ey, double charge,
                                    orc_particle *p, double pos[2], double tStep,
                                    uint64_t seed, uint32_t pid, uint32_t step,
                                    int *contact);
int orc_device_run_philox(const orc_device_run_cfg *cfg, double Vt, double *eqPot,
                          int *nInit, int *nPartOut, double *potOut, int *exitOut,
                          int *delOut, int *injOut, double *nrPartLast, double *stateOut,
                          int maxState);
int orc_device_run_philox_h(orc_device *d, orc_model *m, int pm, double accuracyV,
                            double omega, double tStep, int nSteps, uint64_t seed, double Vt,
                            double *eqPot, int *nInit, int *nPartOut, double *potOut,
                            int *exitOut, int *delOut, int *injOut, double *nrPartLast,
                            double *stateOut, int maxState, int poissonInterval,
                            double chanX0, double chanX1, double chanLen, double *chanOut,
                            int initFromPotential);

/* ---- handle API for language bindings (tests) ---------------------------- */
orc_device *orc_device_new(double maxX, double maxY, double hx, double hy,
                           double T, double epsR, double ni);
void orc_device_delete(orc_device *d);
/* ext[2]; scal = Vt, debye length, cell volume, normalised cell volume */
void orc_device_info(const orc_device *d, int *ext, double *scal);
void orc_device_get_grids(const orc_device *d, double *doping, int *region,
                          int *ohmic);
/* SOR solve in place (conc == NULL: equilibrium); returns sweeps */
long long orc_sor_solve(const orc_device *d, double *pot, const double *conc,
                        double accuracyV, double omega, int resetBC);
void orc_pm_assign_h(const orc_device *d, int scheme, const double *pos, int n,
                     double nrCarriers, double *grid);
void orc_pm_forces_h(const orc_device *d, int scheme, const double *ex,
                     const double *ey, const double *pos, int n, double charge,
                     double *forces);
void orc_device_init_replay(const orc_model *m, const orc_device *d, int np,
                            const int *coords, const uint64_t *draws,
                            size_t stride, size_t *cursor, double *state);
void orc_device_run_replay(const orc_model *m, const orc_device *d, int pm,
                           const double *ex, const double *ey, int np,
                           double *state, const uint64_t *draws, size_t stride,
                           size_t *cursor, int nsteps, double tStep,
                           int *removedStep, int *contact, double *posOut,
                           double *kOut, double *EOut, int32_t *nevOut);
int orc_device_run_h(orc_device *d, orc_model *m, int pm, double accuracyV,
                     double omega, double tStep, int nSteps, uint64_t seed,
                     double *eqPot, int *nInit, int *nPartOut, double *potOut,
                     int *exitOut, int *delOut, int *injOut, double *nrPartLast,
                     double *stateOut, int maxState, int poissonInterval, double chanX0,
                     double chanX1, double chanLen, double *chanOut, int initFromPotential);

#ifdef __cplusplus
}
#endif
#endif /* EMC_ORACLE_DEVICE_H */
