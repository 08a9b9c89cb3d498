// emcgpu_device.h -- parameter blocks shared by the host library and the
// gfx950 kernels of the EMC particle step.
//
// Kinematics on the device run in scaled units so that FP32 neither
// underflows nor overflows (hbar^2 k^2 is ~1e-50 in SI):
//   k in 1/nm, energy in eV, time in fs, velocity in m/s.
// Particle state lives in HBM as structure-of-arrays over an arena holding
// every instance, each instance padded to a whole number of 256-particle
// tiles so that a tile never straddles two instances.
#ifndef EMCGPU_DEVICE_H
#define EMCGPU_DEVICE_H

#include <stdint.h>

#include "../../include/emcgpu.h"

namespace emcgpu {

constexpr int kMaxV = EMCGPU_MAX_VALLEYS;
constexpr int kMaxDeg = EMCGPU_MAX_DEG;
constexpr int kMaxZ = EMCGPU_MAX_Z;
constexpr int kMaxMech = EMCGPU_MAX_MECH;
constexpr int kTile = 256;  // threads per block
// arena unit: every instance starts on a multiple of kUnit particles, so a
// kernel tile of 256 * PPT particles (PPT in {1, 2, 4, 8}) never straddles
// two instances
constexpr int kUnit = 2048;
constexpr int kMaxPPT = 8;
constexpr int kMaxBaths = EMCGPU_MAX_BATHS;
constexpr int kMaxBins = EMCGPU_MAX_BINS;
// copies of the global bath tallies: block b adds its counts to copy
// b % kTallyCopies, so the few hot bins are not hammered by every block
constexpr int kTallyCopies = 64;

// final-state sampler of a mechanism
enum FinalKind : int {
  kFinalIsotropicElastic = 0,  // acoustic: |k| kept, random direction
  kFinalIsotropicInelastic = 1,  // IV: E += shift, new valley/subvalley
  kFinalFroehlich = 2            // polar optical: angle wrt old k
};

template <typename R>
struct ValleyDev {
  R c2;     // gamma(E) = c2 * |k|^2   [eV nm^2]
  R alpha;  // non-parabolicity [1/eV] (0 for parabolic)
  R cv;     // hbar * 1e9 / m_c        [m/s per 1/nm]
  int deg;
  int pad;
  R accel[kMaxDeg][3];  // dk/dt in device frame [1/nm per fs]
  R wdir[kMaxDeg][3];   // v.dir = cv / npf * dot(k, wdir[sub])
};

template <typename R>
struct MechDev {
  int final_kind;
  int final_valley;
  int z;
  int bath1;  // hot-phonon Froehlich: instance-local bath index + 1 (0: none)
  R shift;  // energy change of the final state [eV] (IV, Froehlich)
  uint8_t final_sub[kMaxDeg][kMaxZ];
};

template <typename R>
struct InstDev {
  int n_valleys;
  int n_levels;
  R dE;     // table energy step [eV]
  R inv_dE;
  R dt;     // step [fs]
  R vt;     // thermal voltage [V] for initialisation
  R tau0[kMaxV];  // [fs]
  int n_tables[kMaxV];
  long long table_off[kMaxV];  // element offset of the valley's [level][table] block
  ValleyDev<R> valley[kMaxV];
  MechDev<R> mech[kMaxV][kMaxMech];  // per valley, in table order
  // phonon baths of the instance: their bins are contiguous, bath b at
  // local bin offset bath_loff[b] of the instance's block that starts at
  // global bin gbin_off
  int n_baths;
  int total_bins;
  int bath_nb[kMaxBaths];
  int bath_loff[kMaxBaths];
  R bath_dq[kMaxBaths];  // bin width [1/nm]
  long long gbin_off;
};

// constants of one LO phonon bath (emcPhononBath<double>)
struct BathDev {
  int n_bins, has_ac, has_ridley, has_profile;
  long long off;  // first bin in the global bin arrays
  double dq, tau_lo, n0, vsim, lo_e, lat_t;
  double sum_w;   // sum of q_i^2 over the bins in bin order (getMeanNq weight)
  double ac_e, tau_ac, n_ac_eq;
  double w_r, to_e, tau_to, n_to_eq;
};
// evolving scalars of a bath
struct BathVar {
  double n_ac, n_to, mean, pad;
};

// one (instance, valley) table set rebuilt on the GPU from the baths
constexpr int kRateConst = 0, kRateHotAb = 1, kRateHotEm = 2;
struct RebuildJob {
  int inst, valley, n_tables, n_levels;
  long long cum_off;                // element offset in the [level][table] tables
  int kind[kMaxMech];               // kRate*
  int bath[kMaxMech];               // global bath index (hot kinds)
  double C[kMaxMech];               // rate prefactor (hot kinds)
  long long rate_off[kMaxMech];     // const: rates[level]; hot: kI[level], ln[level]
};

// the baths and rebuild jobs of one instance (one block of k_bath_step)
struct BathInstDev {
  int bath0, n_baths, job0, n_jobs;
};

template <typename R>
struct BathStepArgs {
  const BathInstDev *groups;
  const BathDev *baths;
  double *nq;
  const double *tau_prof;
  unsigned long long *tally;  // [kTallyCopies][tally_bins][2]
  long long tally_bins;
  BathVar *var;
  double dt;
  double *hist;  // [bath][hist_steps] mean occupation after each step, or null
  int hist_steps, hist_step;
  int rebuild;
  const RebuildJob *jobs;
  const double *arena;
  R *cum;
  double *cum64;
  InstDev<R> *inst;
  double *tau0_64;
};

template <typename R>
struct StepArgs {
  R *kx, *ky, *kz, *E, *tau;
  uint32_t *vs;  // valley << 16 | subvalley
  const InstDev<R> *inst;
  const R *cum;
  const int *unit_inst;  // instance of each kUnit-particle arena unit
  const long long *inst_start;
  const long long *inst_count;
  int n_tiles;
  int tile_base;  // first tile of the launch (initialisation of one instance)
  int n_steps;   // steps fused in this launch
  uint32_t step0;
  uint32_t key0, key1;
  double *partial;  // [n_steps][n_obs][n_tiles]
  int n_obs;
  int max_v;
  int max_deg;
  // replay
  const uint64_t *draws;
  long long draw_stride;
  uint64_t *cursor;
  double *grain_tau;  // seconds
  double dt_s;
  int *overflow;
  // trace (instance 0)
  int32_t *tr_valley, *tr_sub, *tr_events;
  double *tr_k, *tr_E;
  long long tr_n;
  // phonon-bath event tallies [kTallyCopies][global bin][2] (emission,
  // absorption), tally_bins bins per copy
  unsigned long long *tally;
  long long tally_bins;
  int lds_bins;  // bins held in LDS per block (max over instances)
};

}  // namespace emcgpu

#endif
