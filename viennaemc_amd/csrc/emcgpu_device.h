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
#include "../include/emcgpu/emcgpuPhysics.hpp"

namespace emcgpu {

constexpr int kMaxV = EMCGPU_MAX_VALLEYS;
constexpr int kMaxDeg = EMCGPU_MAX_DEG;
constexpr int kMaxZ = EMCGPU_MAX_Z;
constexpr int kMaxMech = EMCGPU_MAX_MECH;
constexpr int kTile = 256;  // threads per block
constexpr int kSubs = kTile / 64;  // waves per block: one step partial each
// arena unit: every instance starts on a multiple of kUnit particles, so a
// kernel tile of 256 * PPT particles (PPT in {1, 2, 4, 8}) never straddles
// two instances
constexpr int kUnit = 2048;
constexpr int kMaxPPT = 8;
constexpr int kMaxBaths = EMCGPU_MAX_BATHS;
constexpr int kMaxBins = EMCGPU_MAX_BINS;
// copies of the global bath tallies: block b adds its counts to copy
// b % kTallyCopies.  Workgroups are dealt to the 8 XCDs round-robin, so each
// copy is written from one XCD and its atomics stay in that XCD's L2, and
// the few hot bins are not hammered by every block
constexpr int kTallyCopies = 8;

// final-state sampler of a mechanism
enum FinalKind : int {
  kFinalIsotropicElastic = 0,  // acoustic: |k| kept, random direction
  kFinalIsotropicInelastic = 1,  // IV: E += shift, new valley/subvalley
  kFinalFroehlich = 2,           // polar optical: angle wrt old k
  kFinalScreenedFroehlich = 3    // screened polar optical: angle wrt old k
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
  // Herring-Vogt metric of each subvalley in the device frame,
  // M = R^T diag(vogt) R, symmetric: M00 M01 M02 M11 M12 M22
  R hv[kMaxDeg][6];
};

template <typename R>
struct MechDev {
  int final_kind;
  int final_valley;
  int z;
  int bath1;  // hot-phonon Froehlich: instance-local bath index + 1 (0: none)
  int qres;   // screened hot-phonon: angle through the bath's q distribution
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
  R qs2;  // plasmon screening q_s^2 [1/nm^2]
  // window sums of each bath (cumW[nb + 1] then cumWN[nb + 1], q in 1/m);
  // the instance's baths are contiguous from bath_win[0], 2 (total_bins +
  // n_baths) values
  const R *bath_win[kMaxBaths];
  int win_lds;  // a q-resolved angle reads them: the step kernel stages them in LDS
  // bulk runs with positions: the periodic box [0, box] per dimension [nm]
  // (the device's maxPos, basicBulkParticleHandler.hpp:600-613)
  R box[3];
};

// constants of one LO phonon bath (emcPhononBath<double>)
struct BathDev {
  int n_bins, has_ac, has_ridley, has_profile;
  long long off;  // first bin in the global bin arrays
  double dq, tau_lo, n0, vsim, lo_e, lat_t;
  double sum_w;   // sum of q_i^2 over the bins in bin order (getMeanNq weight)
  double ac_e, tau_ac, n_ac_eq;
  double w_r, to_e, tau_to, n_to_eq;
  long long woff;  // first element of the bath's window sums
};
// evolving scalars of a bath
struct BathVar {
  double n_ac, n_to, mean, pad;
};

// one (instance, valley) table set rebuilt on the GPU from the baths
constexpr int kRateConst = 0, kRateHotAb = 1, kRateHotEm = 2, kRateScrAb = 3, kRateScrEm = 4;
// occupation of a screened rate: constant, bath mean, bath window mean
constexpr int kNConst = 0, kNMean = 1, kNWindow = 2;
struct RebuildJob {
  int inst, valley, n_tables, n_levels;
  long long cum_off;                // element offset in the [level][table] tables
  int kind[kMaxMech];               // kRate*
  int bath[kMaxMech];               // global bath index (hot kinds)
  double C[kMaxMech];               // rate prefactor (hot and screened kinds)
  long long rate_off[kMaxMech];     // const: rates[level]; hot: kI[level], ln[level];
                                    // screened: kI[level], kF[level] (1/m)
  int nsrc[kMaxMech];               // screened: kN*
  double nconst[kMaxMech];          // screened, kNConst: Bose occupation
  int every_step;                   // rebuilt after every step (plasma screening)
};
// one mechanism's raw rates on the table grid, evaluated on the GPU from its
// descriptor (kind: emcgpu_mech_kind of a static rate)
struct RateDescJob {
  int kind, absorption;
  double C, nq, hw, delta, meff, qs2;
  phys::ValleyMass iv, fv;
  long long out_off;  // first level of the mechanism in the output
};

// the per-step rebuild over many blocks: kRebuildChunk levels per block
constexpr int kRebuildChunk = 256;
constexpr int kRebuildChunks = (EMCGPU_MAX_LEVELS + kRebuildChunk - 1) / kRebuildChunk;

// the baths and rebuild jobs of one instance (one block of k_bath_step)
struct BathInstDev {
  int bath0, n_baths, job0, n_jobs;
  int inst, plasma, n_valleys;
  int need_win;  // a q-resolved mechanism reads the baths' window sums
  double scr_eps, scr_vol; // plasma screening
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
  double *energy;  // [bath][hist_steps][3] energy record of each update, or null
  int hist_steps, hist_step;
  int rebuild;
  const RebuildJob *jobs;
  const double *arena;
  R *cum;
  double *cum64;
  InstDev<R> *inst;
  double *tau0_64;
  // screening and window sums
  double *scr;     // q_s^2 per instance [1/m^2]
  double *win64;   // window sums, FP64
  R *win;          // the same in the precision of the kinematics
  const double *partial;  // step partials [n_obs][n_tiles] of the step just taken
  int n_obs, n_tiles, tile_size;
  int n_jobs;
  double *rb_max;  // [job][kRebuildChunks] block maxima of the cumulative rates
  const long long *inst_start, *inst_count;  // particles of each instance
};

template <typename R>
struct StepArgs {
  R *kx, *ky, *kz, *E, *tau;
  uint32_t *vs;  // valley << 16 | subvalley
  R *px, *py, *pz;  // bulk positions [nm], or null (not tracked)
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
  double *partial;  // [n_steps][n_obs][n_tiles * kSubs] (one per wave)
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
  double *tr_pos;  // positions [nm] of the trace, or null
  long long tr_n;
  // phonon-bath event tallies [kTallyCopies][global bin][2] (emission,
  // absorption), tally_bins bins per copy
  unsigned long long *tally;
  long long tally_bins;
  int lds_bins;  // bins held in LDS per block (max over instances)
  int lds_win;   // window-sum values staged in LDS per block (max over instances)
  int ext_kinds; // some instance uses a screened Froehlich kind
};

// ===========================================================================
// 2-D device simulation
// ===========================================================================
constexpr int kDevThreads = 1024;      // field solver: one workgroup
constexpr int kDevStepThreads = 512;   // contact ranks: 8 waves (segments) per block
constexpr int kDevStepWaves = kDevStepThreads / 64;
constexpr int kDevSegPerCU = 16;       // contact-rank segments per CU (4 measured slower)
// particle step: one block per CU whose waves share the CU's chunks through
// an LDS counter (more blocks per CU, or more particles per lane, measured
// slower: the blocks of one CU then finish unevenly); FP32 16 waves, FP64
// 8 (its registers would spill at 16)
template <typename R>
struct DevStepShape {
  static constexpr int threads = sizeof(R) == 4 ? 1024 : 512;
  static constexpr int waves = threads / 64;
};
constexpr uint32_t kHole = 0xFFFFFFFFu;  // valley/subvalley word of a deleted particle
constexpr int kMaxLevels = 12;
constexpr int kPad = 4 * kDevThreads;  // node padding of the Poisson vectors
constexpr int kMaxContacts = EMCGPU_MAX_CONTACTS;
constexpr int kMaxSlots = 1024;        // ohmic nodes
// dynamic LDS of one workgroup (160 KB per CU on gfx950, less the kernels'
// static arrays)
constexpr int kDevMaxLds = 159 * 1024;
// Philox domains of the device streams (the instance word of the counter)
constexpr uint32_t kDomMove = 0x40000000u;
constexpr uint32_t kDomInit = 0x40000001u;
constexpr uint32_t kDomCount = 0x40000002u;

// committed state of the device ensemble (device memory)
struct DevState {
  long long n_total;    // array length: particles + holes
  long long n_alive;
  long long next_pid;   // creation index of the next particle
  long long step;       // step counter (Philox step word)
  long long hist_base;  // step of history row 0
  long long inj_total;  // injections of the current step (contact scan)
  long long capacity;
  int overflow;
  long long avg_n;      // steps added to the running potential / concentration sums
};

// per-step records [row][...] (row = step - hist_base)
struct DevHist {
  int *exits, *dels, *injs;  // [row][n_contacts]
  long long *n_part;         // [row]
  int *iters;                // [row]
  double *resid, *checksum;  // [row]
  double *chan;              // [row] drift-velocity (Ramo) channel current [A]
  int rows;
};

// particle arrays of the device ensemble (SoA, arena of capacity entries)
template <typename R>
struct DevParticles {
  R *x, *y;  // [nm]
  R *kx, *ky, *kz, *E, *tau;
  uint32_t *vs;   // valley << 16 | subvalley, kHole for a deleted particle
  uint32_t *pid;  // creation index (Philox stream of the particle)
};

// a 3-D grid of the particle-mesh operators: nx x ny x nz nodes, x fastest
// (emcGrid.hpp:241-253), spacing [m], scheme EMCGPU_PM_NGP or _CIC
struct PMGrid3 {
  int nx, ny, nz;
  double hx, hy, hz;
  int pm;
};

// geometry for the particle kernels (positions in nm)
template <typename R>
struct DevGeom {
  int nx, ny, nn, pm;
  R hx, hy;        // [nm]
  R maxx, maxy;    // [nm]
  const int *cnode;  // ohmic contact index per node, -1 elsewhere
  const int *cslot;  // contact slot per node, -1 elsewhere
  const R *acc;      // [node][2] interpolated force / hbar [1/nm per fs]
  // per-region tables: the doping region of each node, whose instance holds
  // the tables of that region (null: instance 0 everywhere)
  const int8_t *region;
};

// one multigrid level: node counts, offset into the level-stacked arrays,
// first entries of its x / y tables
struct MGLevel {
  int nx, ny, off, tx, ty;
};
// per level and dimension, per index: FV coefficients and transfer data
struct DimTab {
  float inv_l, inv_r, w;  // 1/left spacing, 1/right spacing, dual width (fine units)
  float a;                // prolongation: weight of coarse node cI (1: coincident)
  int cI;                 // prolongation: left coarse index
  int f;                  // restriction: fine index of this (coarse) node
  float wl, wr;           // restriction: weights of fine nodes f-1, f+1 (mirror folded)
};

// Poisson + E-field on one workgroup
template <typename R>
struct DevFieldArgs {
  int nx, ny, nn, pm;
  double hfx, hfy, hp;  // h-factors and product of the normalised spacings
  double vt, hxm, hym;  // V_T [V], spacing [m]
  double tol;
  int max_iter, equilibrium, reset_bc, do_poisson, do_efield;
  double *pot;                 // normalised potential
  const double *dopn;          // doping / n_i
  const double *conc;          // electron concentration / n_i (non-equilibrium)
  const double *bc_eq, *bc_neq;  // ohmic node values (eq / with the applied voltage)
  // gate (Robin) terms per node, null without gates: [node][3] = g, g V_eq,
  // g V_neq; the residual gains g V - g phi and the Jacobian's shift g
  const double *gate;
  const uint8_t *dmask;        // ohmic node (Dirichlet) of the fine grid
  const signed char *bcx, *bcy;  // E-field boundary: -2 off the boundary, -1 no contact
  double *ex, *ey;             // [V/m]
  R *acc;                      // [node][2] scaled force / hbar
  double acc_scale;            // charge * 1e-24 / hbar
  // FP64 Newton / conjugate-gradient vectors [nn], FP32 copies of the
  // residual and diagonal shift for the multigrid sweeps
  double *r64, *x64, *p64, *q64, *s64;  // [n_pad]
  float2 *rs0;     // [n_pad] FP32 residual and diagonal shift (0, big at Dirichlet nodes)
  const float *w0; // [n_pad] 1 at active nodes, 0 at Dirichlet nodes and padding
  const float *hw0; // [n_pad] control-volume weights of the CG inner product
  int n_pad;       // nn rounded up to kPad
  const float *cinv;    // [n][n] inverse of the direct_level operator (row-major)
  int direct_level;     // coarse level solved exactly (0: none)
  int max_vcycles;             // V-cycles per solve
  int max_pcg;                 // conjugate-gradient iterations per Newton step
  int n_levels;
  MGLevel lev[kMaxLevels];
  const DimTab *tab;
  int n_tab;
  const uint8_t *lmask;        // Dirichlet mask of all levels (level-stacked)
  int n_coarse;                // nodes of the levels below the fine grid
  int wave_level;              // first level run by one wave
  int nu_pre, nu_post, nu_coarse;
  // dynamic LDS layout (float offsets; tab and cmask in bytes)
  int lds_res, lds_rc, lds_sc, lds_tab, lds_cmask;
  int lds_dummy;  // float offset of one scratch word per thread
  // the same layout in global memory (L2-resident) when it exceeds the
  // LDS, null when it fits
  float *mg_ws;
  // fine-grid thread map: a thread owns column (tid mod 2^cw_log2) of a row
  // block; half-row map (red-black sweeps, node pairs) and full-row map,
  // with the number of row blocks (passes) covering the grid
  int hcw_log2, nph, fcw_log2, npf;
  // records
  int *iters_out;
  double *resid_out;
  const double *nrpart;        // check sum input (may be null)
  DevState *state;
  DevHist hist;
  int record;                  // write iters / resid / check sum to the history
  int interval;                // recorded steps: solve only when step % interval == 0
  // dense Newton corrections before this launch: green_iters of them at
  // most, their residual checks in green_acc (see GreenArgs); the launch
  // keeps the solve when the last check met the stopping rule, and clears
  // the checks
  int green_iters;
  // 1: the E-field of a solve kept from the corrections was written by
  // their last check; computed here only when this launch solves
  int efield_if_solved;
  const unsigned long long *green_flag;
  unsigned long long *green_acc;
};

// dense Newton corrections of a per-step Poisson solve: the inverse of the
// control-volume weighted operator over the active nodes, G = (H A0)^-1
// (A0: the discrete operator of the Newton step without the hole term
// h^2 p, gate shifts included, H: the control-volume weights), applied as
// x = G H F
constexpr int kGreenThreads = 1024;  // apply kernel: 16 waves, one row per wave at a time
constexpr int kGreenMaxActive = 24576;

// separable devices: the active nodes are the tensor product of mx columns
// and my rows (Dirichlet nodes fill whole columns or rows) and the gate
// shifts split into a column and a row part, so A0 = Ax (+) Ay (Kronecker
// sum of the 1-D operators over the active columns and rows) and
// A0^-1 = (Bx (x) By) (Lx (+) Ly)^-1 (Px (x) Py), with Ax = Bx Lx Px the
// eigen-decomposition of each 1-D operator (Px = Bx^-1).  FP64 matrices;
// the x ones stored transposed so a thread per column reads them coalesced.
constexpr int kFdThreads = 512;
struct FdOps {
  int mx, my;
  const int *ix, *iy;        // node column / row of each active column / row
  const double *py;          // [my][my] Py[j][b]
  const double *pxt;         // [mx][mx] PxT[a][k] = Px[k][a]
  const double *by;          // [my][my] By[j][l]
  const double *bxt;         // [mx][mx] BxT[k][a] = Bx[a][k]
  const double *inv_den;     // [my][mx] 1 / (lx[k] + ly[j])
  double *t;                 // [my][mx] spectral coefficients between the two launches
};

struct GreenArgs {
  int nx, ny, nn, ld;          // grid; row stride of G (multiple of 256)
  double hfx, hfy, hp;
  double *pot;                 // normalised potential (updated at active nodes)
  const double *conc, *dopn;   // electron concentration, doping / n_i
  const double *gate;          // gate terms [node][3] (DevFieldArgs::gate), or null
  const float *w0, *hw0;       // active-node mask, control-volume weights
  const float *green;          // [nn][ld], rows and columns of Dirichlet nodes zero
  double *g;                   // [nn] H F of the current potential
  // step + 1 of a step whose hole term is too large for the inverse: its
  // corrections are skipped (the multigrid path then solves alone)
  unsigned long long *flag;
  double shift_max;            // largest h^2 p the inverse may ignore
  // residual check c (before correction c; c = iters: after the last) as
  // bit patterns of non-negative doubles, [iters + 1][2]: max |F| and the
  // scale max h^2 (|C| + n + p) of the stopping rule |F| <= tol * scale
  unsigned long long *acc;
  double tol;
  const DevState *state;
  int interval;                // solve only when step % interval == 0
  // 1: corrections by the separable inverse (fd; g holds F), 0: by G (g = H F)
  int use_fd;
  FdOps fd;
  // with fuse_efield, the last check (c == iters) also writes the E-field
  // and the force grid of the corrected potential (the formulas of
  // k_dev_efield); the field kernel recomputes them only when it changes
  // the potential itself
  int fuse_efield, iters, pm, acc_fp64;
  double vt, hxm, hym, acc_scale;
  const signed char *bcx, *bcy;
  double *ex, *ey;
  void *force;  // the force grid (DevFieldArgs::acc), float or double [node][2] (acc_fp64)
};

// particle step of a device ensemble
template <typename R>
struct DevStepArgs {
  DevParticles<R> p;
  DevGeom<R> g;
  const InstDev<R> *inst;
  const R *cum;
  DevState *state;
  DevHist hist;
  uint32_t key0, key1;
  int n_contacts, n_slots;
  R *dep;                 // [block][node] deposits of the step kernel
  uint32_t *seg_cnt;      // [slot][segment] contact particles per segment
  uint32_t *seg_nst;      // [segment] staged entries
  uint2 *stage;           // (particle, slot << 22 | rank in segment) at segment start
  int16_t *pslot;         // [particle] ohmic slot of the position after the move, -1 if none
  int *exit_part;         // [block][contact] particles that left through each contact
  int lds_dep;            // deposit grid in LDS (else straight to dep row 0 with atomics)
  int move;              // 0: classify positions only (contact rule of given positions)
  // drift-velocity channel current: sum of v_x over the moved particles
  // with chan_x0 <= x <= chan_x1 [nm], one FP64 partial per block
  int chan;
  R chan_x0, chan_x1;
  double *chan_part;      // [block]
  // frozen-field drift (emcgpu_device_drift)
  long long n;            // particles (drift)
  uint32_t step;          // first Philox step (drift)
  int n_steps;
  int *exit_step, *exit_contact;
  double *tr_pos, *tr_k, *tr_E;
  int *tr_events;
  const uint64_t *draws;
  long long draw_stride;
  uint64_t *cursor;
  double *grain_tau;
  double dt_s;
  int *draw_overflow;
};

// contact scan, resolve (deletions + injections), reduce
template <typename R>
struct DevContactArgs {
  DevParticles<R> p;
  DevGeom<R> g;
  const InstDev<R> *inst;
  DevState *state;
  DevHist hist;
  uint32_t key0, key1;
  int n_contacts, n_slots, n_segments, n_blocks_dep;
  const uint32_t *seg_cnt;  // [slot][segment]
  uint32_t *seg_off;       // [slot][segment] exclusive offsets
  const uint32_t *seg_nst;
  const uint2 *stage;
  const int *slot_node;    // grid node of each slot
  const int *slot_contact;
  const double *slot_exp;  // expected carriers at the node
  int *slot_kept, *slot_inj, *slot_total;
  const int *exit_part;    // [step block][contact] exits of the step kernel's blocks
  int n_step_blocks;
  double *corr;            // [node] deposit corrections (FP64 atomics)
  R *dep;
  double *dpart;           // [n_chunks][node] partial sums of the block grids
  int rows_per_chunk, n_chunks;
  double *nrpart, *conc;
  const double *cfac;      // conc = nrpart * cfac
  R vt;                    // thermal voltage [V]
  // channel current: block partials of the step kernel, FP64 corrections of
  // the deleted / injected particles, scale q * carriers / length [A s/m]
  int chan;
  R chan_x0, chan_x1;
  const double *chan_part;
  int n_chan_blocks;
  double *chan_corr;
  double chan_scale;
  // running sums of the potential and concentration (final averages)
  int accumulate;
  const double *pot;
  double *avg_pot, *avg_conc;
};

}  // namespace emcgpu

#endif
