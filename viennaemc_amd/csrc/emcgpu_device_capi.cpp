// emcgpu_device_capi.cpp -- C ABI of the 2-D device simulation (host side).
//
// Builds the device geometry on the host as emcDevice does (extent, doping
// regions, ohmic contacts and their occurrences on every boundary they touch,
// emcDevice.hpp:83-343, emcDopingProfile.hpp:59-207, emcSurface.hpp:205-408),
// lowers it to per-node arrays and a multigrid hierarchy for the Poisson
// kernel, owns the device ensemble (two SoA buffer sets for the stable
// compaction) and drives the per-step kernel sequence of
// emcgpu_device_kernels.hip in chunks, captured in a HIP graph on request.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "emcgpu_host.h"
#include "emcgpu_particle.h"

namespace emcgpu {
template <typename R>
hipError_t dev_launch_field(const DevFieldArgs<R> *args_dev, size_t lds, bool ws, hipStream_t st);
template <typename R>
hipError_t dev_launch_efield(const DevFieldArgs<R> *args_dev, int nn, hipStream_t st);
template <typename R>
hipError_t dev_launch_step(const DevStepArgs<R> &a, const DevContactArgs<R> &ca, int blocks,
                           int n_seg, size_t lds, hipStream_t st);
template <typename R>
hipError_t dev_launch_drift(const DevStepArgs<R> &a, int mode, hipStream_t st);
template <typename R>
hipError_t dev_launch_contacts(const DevContactArgs<R> &a, long long max_inj, hipStream_t st);
template <typename R>
hipError_t dev_launch_reduce(const DevContactArgs<R> &a, int commit, hipStream_t st);
hipError_t dev_launch_green(const GreenArgs &a, int iters, int blocks, hipStream_t st);
hipError_t dev_launch_green_scatter(const int *rows, const int *cols, const double *vals, int n,
                                    double *A, long long lda, hipStream_t st);
hipError_t dev_launch_green_expand(const double *Ainv, long long lda, const int *idx, int nn,
                                   int ld, float *G, hipStream_t st);
hipError_t dev_green_lds(size_t lds);
template <typename R>
hipError_t dev_launch_init(const DevContactArgs<R> &a, const long long *prefix, long long n,
                           hipStream_t st);
template <typename R>
hipError_t dev_launch_compact(const DevParticles<R> &src, const DevParticles<R> &dst, int *blk,
                              long long n_total, DevState *state, hipStream_t st);
template <typename R>
hipError_t dev_launch_zero_r(R *p, long long n, hipStream_t st);
template <typename R>
int dev_step_blocks_per_cu(size_t lds);
hipError_t dev_launch_zero(double *p, long long n, hipStream_t st);
hipError_t dev_launch_zero_i(int *p, long long n, hipStream_t st);
hipError_t dev_launch_mark(DevState *state, hipStream_t st);
hipError_t dev_launch_assign_pos(const DevGeom<double> &g, const double *pos, long long n,
                                 double *grid, hipStream_t st);
hipError_t dev_launch_interp_pos(const DevGeom<double> &g, const double *pos, long long n,
                                 double charge, double *force, hipStream_t st);
hipError_t dev_launch_pm3_assign(const PMGrid3 &g, const double *pos, long long n, double *grid,
                                 hipStream_t st);
hipError_t dev_launch_pm3_interp(const PMGrid3 &g, const double *ex, const double *ey,
                                 const double *ez, const double *pos, long long n, double charge,
                                 double *force, hipStream_t st);
hipError_t dev_launch_pm3_efield(const PMGrid3 &g, double vt, const signed char *bc,
                                 const double *pot, double *ex, double *ey, double *ez,
                                 hipStream_t st);

namespace host {

constexpr double kNm = 1e9;        // m -> nm
constexpr int kChunk = 32;         // steps between host checks of the ensemble
constexpr double kHoleFrac = 0.05; // compaction threshold

struct DeviceSim {
  // ---- geometry (emcDevice) ----------------------------------------------
  int ext[2] = {0, 0};
  int nn = 0;
  double maxPos[2] = {0, 0}, h[2] = {0, 0};
  double width = 1e-6, T = 300, epsR = 1, ni = 1;
  double Vt = 0, debye = 0, cellVolume = 0, normCellVolume = 0;
  std::vector<double> doping;
  std::vector<int> region;
  std::vector<int> contactGrid[4];  // per boundary XMIN, XMAX (ext[1]), YMIN, YMAX (ext[0])
  int nContacts = 0;
  double contactV[kMaxContacts] = {0};
  int contactType[kMaxContacts] = {0};
  double gateEps[kMaxContacts] = {0}, gateThick[kMaxContacts] = {0},
         gateBarrier[kMaxContacts] = {0};
  bool has_gate = false;
  int pm = EMCGPU_PM_CIC;
  bool init_pot = false;  // initial carriers from the equilibrium potential
  bool region_tables = false;  // instance r holds the tables of doping region r
  int n_regions = 0;           // doping regions (instances the tables need)
  double tol = 1e-10;
  int max_iter = 50;
  // ---- per node / per ohmic node ----------------------------------------
  std::vector<int> cnode, cslot;
  std::vector<signed char> bcx, bcy;
  std::vector<uint8_t> dmask;
  std::vector<double> dopn, bc_eq, bc_neq, cfac;
  std::vector<double> gate;  // [node][3] Robin terms (empty without gates)
  std::vector<double> gate_xy;  // [node][2] shift of the x-boundary and y-boundary gates
  std::vector<int> slot_node, slot_contact;
  std::vector<double> slot_exp;
  long long max_inj = 0;  // bound of the injections of one step
  // ---- Poisson ------------------------------------------------------------
  double hfx = 1, hfy = 1, hp = 1;
  std::vector<MGLevel> lev;
  std::vector<DimTab> tab;
  std::vector<uint8_t> lmask;
  size_t mg_lds = 0;
  bool mg_global = false;  // the multigrid data exceed the LDS: global workspace d_mg_ws
  int mg_nodes = 0;  // nodes of all levels
  int lds_res = 0, lds_rc = 0, lds_sc = 0, lds_tab = 0, lds_cmask = 0, wave_level = 0;
  int lds_dummy = 0;
  int direct_level = 0;
  int hcw_log2 = 0, nph = 0, fcw_log2 = 0, npf = 0;
  std::vector<float> cinv_t;
  float *d_cinv_t = nullptr;
  float *d_mg_ws = nullptr;
  // ---- device memory ------------------------------------------------------
  bool on_gpu = false;
  int *d_cnode = nullptr, *d_cslot = nullptr;
  int8_t *d_region = nullptr;
  signed char *d_bcx = nullptr, *d_bcy = nullptr;
  uint8_t *d_dmask = nullptr;
  double *d_dopn = nullptr, *d_bc_eq = nullptr, *d_bc_neq = nullptr, *d_cfac = nullptr;
  double *d_gate = nullptr;
  double *d_pot = nullptr, *d_conc = nullptr, *d_nrpart = nullptr, *d_ex = nullptr;
  double *d_ey = nullptr, *d_corr = nullptr;
  void *d_acc = nullptr;
  double *d_r64 = nullptr, *d_x64 = nullptr, *d_p64 = nullptr, *d_q64 = nullptr;
  double *d_s64 = nullptr;
  float2 *d_rs0 = nullptr;
  float *d_w0 = nullptr, *d_hw0 = nullptr;
  void *d_fargs = nullptr;  // device copies of the field-kernel arguments [slot]
  int n_pad = 0;
  int *d_slot_node = nullptr, *d_slot_contact = nullptr, *d_slot_kept = nullptr;
  int *d_slot_inj = nullptr, *d_slot_total = nullptr;
  int *d_exit_part = nullptr;  // [step block][contact]
  double *d_slot_exp = nullptr;
  DimTab *d_tab = nullptr;
  uint8_t *d_lmask = nullptr;
  DevState *d_state = nullptr;
  int *d_iters = nullptr;
  double *d_resid = nullptr;
  // scratch of the known-answer entry points
  double *d_pot_t = nullptr, *d_conc_t = nullptr, *d_ex_t = nullptr, *d_ey_t = nullptr;
  // ensemble: two buffer sets (stable compaction ping-pong)
  long long capacity = 0;
  int cur = 0;
  void *px[2] = {}, *py[2] = {}, *pkx[2] = {}, *pky[2] = {}, *pkz[2] = {}, *pE[2] = {},
       *ptau[2] = {};
  uint32_t *pvs[2] = {}, *ppid[2] = {};
  uint2 *d_stage = nullptr;
  int16_t *d_pslot = nullptr;
  int *d_blk = nullptr;
  // step kernel
  int blocks = 0, n_seg = 0;
  size_t step_lds = 0;
  bool lds_dep = true;
  void *d_dep = nullptr;
  double *d_dpart = nullptr;
  int rows_per_chunk = 1, n_chunks = 1;
  uint32_t *d_seg_cnt = nullptr, *d_seg_off = nullptr, *d_seg_nst = nullptr;
  // per-step history of a call
  int hist_rows = 0;
  int *d_hexits = nullptr, *d_hdels = nullptr, *d_hinjs = nullptr, *d_hiters = nullptr;
  long long *d_hnpart = nullptr;
  double *d_hresid = nullptr, *d_hcheck = nullptr, *d_hchan = nullptr;
  // options of the current emcgpu_device_step call
  int interval = 1;                 // Poisson solve every interval-th step
  int chan = 0;                     // drift-velocity channel current
  double chan_x0 = 0, chan_x1 = 0;  // [nm]
  double chan_scale = 0;            // q * carriers / length [A s/m]
  int accumulate = 0;               // add potential / concentration to the sums
  double *d_chan_part = nullptr, *d_chan_corr = nullptr;
  double *d_avg_pot = nullptr, *d_avg_conc = nullptr;
  // run state
  bool initialized = false, bc_applied = false;
  uint64_t seed = 0;
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0, graph_cur = -1;
  long long graph_cap = 0;
  // kernel arguments the captured graph was recorded with
  int graph_chan = -1, graph_acc = -1, graph_green = -1, graph_interval = -1;
  double graph_x0 = 0, graph_x1 = 0, graph_scale = 0;
  // dense Newton corrections of the per-step solves: G = (H A0)^-1 over the
  // active nodes as [nn][green_ld] FP32 (0: not built yet, 1: built,
  // -1: not available for this geometry)
  int green_state = 0;
  int green_ld = 0, green_blocks = 0;
  int green_use = 0;  // the current call applies them
  float *d_green = nullptr;
  unsigned long long *d_green_flag = nullptr;
  unsigned long long *d_green_acc = nullptr;  // residual checks of a step
  // separable inverse of the same operator (0: not tried, 1: built, -1: the
  // device is not separable); fd_use: the current call applies it
  int fd_state = 0, fd_use = 0;
  FdOps fd{};
  int *d_fd_ix = nullptr, *d_fd_iy = nullptr;
  double *d_fd_mats = nullptr;  // py, pxt, by, bxt, inv_den, t
};

void dev_free_particles(DeviceSim *d) {
  for (int b = 0; b < 2; b++) {
    dfree(d->px[b]);
    dfree(d->py[b]);
    dfree(d->pkx[b]);
    dfree(d->pky[b]);
    dfree(d->pkz[b]);
    dfree(d->pE[b]);
    dfree(d->ptau[b]);
    dfreet(d->pvs[b]);
    dfreet(d->ppid[b]);
  }
  dfreet(d->d_stage);
  dfreet(d->d_pslot);
  dfreet(d->d_blk);
  d->capacity = 0;
}

void device_sim_destroy(DeviceSim *d) {
  if (!d)
    return;
  if (d->on_gpu) {
    dev_free_particles(d);
    dfreet(d->d_cnode);
    dfreet(d->d_region);
    dfreet(d->d_cslot);
    dfreet(d->d_bcx);
    dfreet(d->d_bcy);
    dfreet(d->d_dmask);
    dfreet(d->d_gate);
    dfreet(d->d_dopn);
    dfreet(d->d_bc_eq);
    dfreet(d->d_bc_neq);
    dfreet(d->d_cfac);
    dfreet(d->d_pot);
    dfreet(d->d_conc);
    dfreet(d->d_nrpart);
    dfreet(d->d_ex);
    dfreet(d->d_ey);
    dfreet(d->d_corr);
    dfree(d->d_acc);
    dfreet(d->d_r64);
    dfreet(d->d_x64);
    dfreet(d->d_p64);
    dfreet(d->d_q64);
    dfreet(d->d_s64);
    dfreet(d->d_rs0);
    dfreet(d->d_w0);
    dfree(d->d_fargs);
    dfreet(d->d_hw0);
    dfreet(d->d_cinv_t);
    dfreet(d->d_mg_ws);
    dfreet(d->d_green);
    dfreet(d->d_green_flag);
    dfreet(d->d_green_acc);
    dfreet(d->d_fd_ix);
    dfreet(d->d_fd_iy);
    dfreet(d->d_fd_mats);
    dfreet(d->d_slot_node);
    dfreet(d->d_slot_contact);
    dfreet(d->d_slot_kept);
    dfreet(d->d_slot_inj);
    dfreet(d->d_slot_total);
    dfreet(d->d_exit_part);
    dfreet(d->d_slot_exp);
    dfreet(d->d_tab);
    dfreet(d->d_lmask);
    dfreet(d->d_state);
    dfreet(d->d_iters);
    dfreet(d->d_resid);
    dfreet(d->d_pot_t);
    dfreet(d->d_conc_t);
    dfreet(d->d_ex_t);
    dfreet(d->d_ey_t);
    dfree(d->d_dep);
    dfreet(d->d_dpart);
    dfreet(d->d_seg_cnt);
    dfreet(d->d_seg_off);
    dfreet(d->d_seg_nst);
    dfreet(d->d_hexits);
    dfreet(d->d_hdels);
    dfreet(d->d_hinjs);
    dfreet(d->d_hiters);
    dfreet(d->d_hnpart);
    dfreet(d->d_hresid);
    dfreet(d->d_hcheck);
    dfreet(d->d_hchan);
    dfreet(d->d_chan_part);
    dfreet(d->d_chan_corr);
    dfreet(d->d_avg_pot);
    dfreet(d->d_avg_conc);
    if (d->gexec)
      (void)hipGraphExecDestroy(d->gexec);
  }
  delete d;
}

namespace {

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
int idx(const DeviceSim &d, int ix, int iy) { return iy * d.ext[0] + ix; }

// posToCoord (emcDevice.hpp:290-297) of one dimension
long long coord_of(double pos, double h) { return (long long)std::round(pos / h); }

bool on_boundary(const DeviceSim &d, int ix, int iy, int b) {
  const int c = b / 2 == 0 ? ix : iy;
  const int dim = b / 2;
  return b % 2 == 1 ? c == d.ext[dim] - 1 : c == 0;
}

// first boundary (XMIN, XMAX, YMIN, YMAX) a node lies on, 4 if none
// (emcSurface.hpp:92-99)
int boundary_of(const DeviceSim &d, int ix, int iy) {
  for (int b = 0; b < 4; b++)
    if (on_boundary(d, ix, iy, b))
      return b;
  return 4;
}

// position of a node along boundary b
int along(int ix, int iy, int b) { return b / 2 == 0 ? iy : ix; }

// contact index of a node, -1 where none (getContactIdx, emcSurface.hpp:251-259)
int contact_of(const DeviceSim &d, int ix, int iy) {
  const int b = boundary_of(d, ix, iy);
  return b == 4 ? -1 : d.contactGrid[b][along(ix, iy, b)];
}

emcgpu_status build_geometry(emcgpu_ctx *c, DeviceSim &d, const emcgpu_device_desc &g) {
  for (int i = 0; i < 2; i++) {
    if (!(g.max_pos[i] > 0) || !(g.spacing[i] > 0))
      return fail(c, EMCGPU_ERR_ARG, "max_pos and spacing must be positive");
    d.maxPos[i] = g.max_pos[i];
    d.h[i] = g.spacing[i];
    // maxPosToExtent (emcUtil.hpp:112-121)
    d.ext[i] = (int)(std::round(g.max_pos[i] / g.spacing[i]) + 1);
    if (d.ext[i] < 3)
      return fail(c, EMCGPU_ERR_ARG, "the device needs at least 3 nodes per dimension");
  }
  if (!(g.temperature > 0) || !(g.eps_r > 0) || !(g.ni > 0) || !(g.width > 0))
    return fail(c, EMCGPU_ERR_ARG, "temperature, eps_r, ni and width must be positive");
  if (g.pm_scheme < EMCGPU_PM_NGP || g.pm_scheme > EMCGPU_PM_NEC)
    return fail(c, EMCGPU_ERR_ARG, "unknown particle-mesh scheme %d", g.pm_scheme);
  d.nn = d.ext[0] * d.ext[1];
  d.T = g.temperature;
  d.epsR = g.eps_r;
  d.ni = g.ni;
  d.width = g.width;
  d.pm = g.pm_scheme;
  d.tol = g.poisson_tol > 0 ? g.poisson_tol : 1e-10;
  d.max_iter = g.poisson_max_iter > 0 ? g.poisson_max_iter : 50;
  d.init_pot = g.init_from_potential != 0;
  d.region_tables = g.region_tables != 0;
  // emcDevice ctor (emcDevice.hpp:83-99): thermal voltage, intrinsic Debye
  // length; calcCellVolume (:333-343)
  d.Vt = phys::kB / phys::q * d.T;
  d.debye = std::sqrt(phys::eps0 * d.epsR * d.Vt / phys::q / d.ni);
  d.cellVolume = 1.0;
  d.cellVolume = d.cellVolume * d.h[0];
  d.cellVolume = d.cellVolume * d.h[1];
  d.normCellVolume = 1.0;
  d.normCellVolume = d.normCellVolume * (d.h[0] / d.debye);
  d.normCellVolume = d.normCellVolume * (d.h[1] / d.debye);
  d.cellVolume *= d.width;
  d.normCellVolume *= d.width / d.debye;
  // doping: n_i everywhere (emcDopingProfile.hpp:59-63), then the regions
  // in order on their node boxes (:74-81, :161-207)
  d.doping.assign(d.nn, d.ni);
  d.region.assign(d.nn, -1);
  if (g.n_doping < 0 || g.n_doping > EMCGPU_MAX_REGIONS || (g.n_doping > 0 && !g.doping))
    return fail(c, EMCGPU_ERR_ARG, "n_doping must be in [0, %d]", EMCGPU_MAX_REGIONS);
  for (int r = 0; r < g.n_doping; r++) {
    const emcgpu_doping_region &dr = g.doping[r];
    long long lo[2], hi[2];
    for (int i = 0; i < 2; i++) {
      long long a = coord_of(dr.min_pos[i], d.h[i]), b = coord_of(dr.max_pos[i], d.h[i]);
      if (a < 0 || b < 0)
        return fail(c, EMCGPU_ERR_ARG, "doping region %d has a negative corner", r);
      if (b < a)
        std::swap(a, b);
      if (a > d.ext[i])
        return fail(c, EMCGPU_ERR_ARG, "doping region %d lies outside the device", r);
      lo[i] = a;
      hi[i] = std::min<long long>(b, d.ext[i]);
    }
    for (int iy = 0; iy < d.ext[1]; iy++)
      for (int ix = 0; ix < d.ext[0]; ix++)
        if (ix >= lo[0] && ix <= hi[0] && iy >= lo[1] && iy <= hi[1]) {
          d.doping[idx(d, ix, iy)] = dr.doping;
          d.region[idx(d, ix, iy)] = r;
        }
  }
  d.n_regions = g.n_doping;
  if (d.region_tables) {
    if (g.n_doping > c->n_inst || g.n_doping > 127)
      return fail(c, EMCGPU_ERR_ARG,
                  "per-region tables need an instance per doping region (%d regions, %d instances)",
                  g.n_doping, c->n_inst);
    for (int k = 0; k < d.nn; k++)
      if (d.region[k] < 0)
        return fail(c, EMCGPU_ERR_ARG, "per-region tables: node %d lies in no doping region", k);
  }
  // contacts (addContact, emcSurface.hpp:199-222): an ohmic contact is
  // marked on every boundary its nodes lie on (updateAllOccurences,
  // :381-408), a gate on its own boundary only
  for (int b = 0; b < 4; b++)
    d.contactGrid[b].assign(d.ext[b < 2 ? 1 : 0], -1);
  if (g.n_contacts < 0 || g.n_contacts > EMCGPU_MAX_CONTACTS ||
      (g.n_contacts > 0 && !g.contacts))
    return fail(c, EMCGPU_ERR_ARG, "n_contacts must be in [0, %d]", EMCGPU_MAX_CONTACTS);
  d.nContacts = g.n_contacts;
  for (int k = 0; k < g.n_contacts; k++) {
    const emcgpu_contact &ct = g.contacts[k];
    const int b = ct.boundary;
    if (b < 0 || b > 3)
      return fail(c, EMCGPU_ERR_ARG, "contact %d: unknown boundary %d", k, b);
    const double hs = b / 2 == 0 ? d.h[1] : d.h[0];
    const long long mn = coord_of(ct.min_pos, hs), mx = coord_of(ct.max_pos, hs);
    const long long len = (long long)d.contactGrid[b].size();
    if (mn < 0 || mx < mn || mx >= len)
      return fail(c, EMCGPU_ERR_ARG, "contact %d lies outside its boundary", k);
    if (ct.type != EMCGPU_CONTACT_OHMIC && ct.type != EMCGPU_CONTACT_GATE)
      return fail(c, EMCGPU_ERR_ARG,
                  ct.type == EMCGPU_CONTACT_SCHOTTKY
                      ? "contact %d: Schottky contacts are not supported"
                      : "contact %d: unknown contact type",
                  k);
    const bool gate = ct.type == EMCGPU_CONTACT_GATE;
    if (gate && !(ct.eps_oxide > 0 && ct.thickness > 0))
      return fail(c, EMCGPU_ERR_ARG, "gate contact %d needs a positive eps_oxide and thickness",
                  k);
    for (long long s = mn; s <= mx; s++) {
      const int ix = b == 0 ? 0 : b == 1 ? d.ext[0] - 1 : (int)s;
      const int iy = b == 2 ? 0 : b == 3 ? d.ext[1] - 1 : (int)s;
      for (int bb = 0; bb < 4; bb++)
        if (bb == b || (!gate && on_boundary(d, ix, iy, bb)))
          d.contactGrid[bb][along(ix, iy, bb)] = k;
    }
    d.contactV[k] = ct.voltage;
    d.contactType[k] = ct.type;
    d.gateEps[k] = gate ? ct.eps_oxide : 0.0;
    d.gateThick[k] = gate ? ct.thickness : 0.0;
    d.gateBarrier[k] = gate ? ct.barrier_height : 0.0;
    d.has_gate = d.has_gate || gate;
  }
  return EMCGPU_OK;
}

// per-node arrays of the kernels, ohmic-node slots, SOR constants
// (initDeviceCharacteristics, emcSORSolver.hpp:330-366)
void derive_nodes(DeviceSim &d) {
  const int nx = d.ext[0], ny = d.ext[1];
  const double hn0 = d.h[0] / d.debye, hn1 = d.h[1] / d.debye;
  d.hfx = hn1 / hn0;
  d.hfy = hn0 / hn1;
  d.hp = 1.0;
  d.hp = d.hp * hn0;
  d.hp = d.hp * hn1;
  d.gate.clear();
  d.gate_xy.clear();
  if (d.has_gate) {
    d.gate.assign((size_t)3 * d.nn, 0.0);
    d.gate_xy.assign((size_t)2 * d.nn, 0.0);
  }
  d.cnode.assign(d.nn, -1);
  d.cslot.assign(d.nn, -1);
  d.bcx.assign(d.nn, -2);
  d.bcy.assign(d.nn, -2);
  d.dmask.assign(d.nn, 0);
  d.dopn.resize(d.nn);
  d.bc_eq.assign(d.nn, 0.0);
  d.bc_neq.assign(d.nn, 0.0);
  d.cfac.resize(d.nn);
  d.slot_node.clear();
  d.slot_contact.clear();
  d.slot_exp.clear();
  d.max_inj = 0;
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      const int k = idx(d, ix, iy);
      d.dopn[k] = d.doping[k] / d.ni;
      // isOhmicContact (emcSurface.hpp:240-246): the contact of the node's
      // first boundary, if ohmic
      int ci = contact_of(d, ix, iy);
      if (ci >= 0 && d.contactType[ci] != EMCGPU_CONTACT_OHMIC)
        ci = -1;
      d.cnode[k] = ci;
      if (d.has_gate)
        for (int b = 0; b < 4; b++) {
          // checkForGateContact (emcSORSolver.hpp:399-412): a gate of each
          // boundary the node lies on adds g hF_d h_d (V_g - phi) to the
          // residual, g = 2 (eps_ox / eps_r) / (t_ox / L_D), V_g the barrier
          // height (equilibrium) plus the applied voltage (non-equilibrium)
          if (!on_boundary(d, ix, iy, b))
            continue;
          const int gi = d.contactGrid[b][along(ix, iy, b)];
          if (gi < 0 || d.contactType[gi] != EMCGPU_CONTACT_GATE)
            continue;
          const double gf = 2.0 * (d.gateEps[gi] / d.epsR) / (d.gateThick[gi] / d.debye);
          const double hh = b / 2 == 0 ? d.hfx * hn0 : d.hfy * hn1;
          const double vb = d.gateBarrier[gi] / d.Vt, va = d.contactV[gi] / d.Vt;
          d.gate[3 * (size_t)k] += gf * hh;
          d.gate_xy[2 * (size_t)k + b / 2] += gf * hh;
          d.gate[3 * (size_t)k + 1] += gf * vb * hh;
          d.gate[3 * (size_t)k + 2] += gf * (vb + va) * hh;
        }
      if (ix == 0)
        d.bcx[k] = (signed char)d.contactGrid[0][iy];
      else if (ix == nx - 1)
        d.bcx[k] = (signed char)d.contactGrid[1][iy];
      if (iy == 0)
        d.bcy[k] = (signed char)d.contactGrid[2][ix];
      else if (iy == ny - 1)
        d.bcy[k] = (signed char)d.contactGrid[3][ix];
      int nb = 0;
      if (ix == 0 || ix == nx - 1)
        nb++;
      if (iy == 0 || iy == ny - 1)
        nb++;
      // updateCurrentParticleConcentrations (emcSimulationResults.hpp:98-116)
      d.cfac[k] = (1.0 / d.ni) * (1.0 / d.cellVolume) * (nb == 2 ? 4.0 : nb == 1 ? 2.0 : 1.0);
      if (ci >= 0) {
        d.dmask[k] = 1;
        // ohmic boundary values (emcSORSolver.hpp:142-145, :288-292)
        const double bi = std::asinh(0.5 * d.dopn[k]);
        d.bc_eq[k] = bi;
        d.bc_neq[k] = d.contactV[ci] / d.Vt + bi;
        // getExpectedNrParticlesAtContact (emcElectron.hpp:63-73)
        double e = d.cellVolume * d.doping[k];
        if (ix == 0 || ix == nx - 1)
          e *= 0.5;
        if (iy == 0 || iy == ny - 1)
          e *= 0.5;
        d.cslot[k] = (int)d.slot_node.size();
        d.slot_node.push_back(k);
        d.slot_contact.push_back(ci);
        d.slot_exp.push_back(e);
        d.max_inj += (long long)std::ceil(e);
      }
    }
}

// coarse positions: every other node; when that misses the last node, the
// last node is appended if the leftover interval is at least half the
// stride and otherwise replaces the last kept node, so coarse intervals stay
// within 0.5 - 1.5 strides; dimensions of 3 or fewer nodes stay
std::vector<int> coarsen(const std::vector<int> &p) {
  if (p.size() <= 3)
    return p;
  std::vector<int> q;
  for (size_t i = 0; i < p.size(); i += 2)
    q.push_back(p[i]);
  if (q.back() != p.back()) {
    const int left = p.back() - q.back(), stride = q.back() - q[q.size() - 2];
    if (2 * left >= stride)
      q.push_back(p.back());
    else
      q.back() = p.back();
  }
  return q;
}

// transfer and coefficient tables of one dimension of level l
void dim_tables(const std::vector<int> &p, const std::vector<int> *coarse,
                const std::vector<int> *fine, std::vector<DimTab> &out) {
  const int n = (int)p.size();
  for (int i = 0; i < n; i++) {
    DimTab t{};
    const double dl = i > 0 ? p[i] - p[i - 1] : p[1] - p[0];
    const double dr = i < n - 1 ? p[i + 1] - p[i] : p[n - 1] - p[n - 2];
    t.inv_l = (float)(1.0 / dl);
    t.inv_r = (float)(1.0 / dr);
    t.w = (float)(0.5 * (dl + dr));
    // prolongation from the coarser level: left coarse node and its weight
    t.a = 1.0f;
    t.cI = 0;
    if (coarse) {
      const std::vector<int> &q = *coarse;
      int I = (int)(std::upper_bound(q.begin(), q.end(), p[i]) - q.begin()) - 1;
      t.cI = I;
      if (q[I] != p[i])
        t.a = (float)((double)(q[I + 1] - p[i]) / (double)(q[I + 1] - q[I]));
    }
    // restriction from the finer level: this node's fine index and the
    // weights of its fine neighbours (mirror images folded in)
    t.f = i;
    t.wl = t.wr = 0.0f;
    if (fine) {
      const std::vector<int> &f = *fine;
      const int fi = (int)(std::lower_bound(f.begin(), f.end(), p[i]) - f.begin());
      t.f = fi;
      double wl = 0.0, wr = 0.0;
      if (fi > 0 && i > 0 && f[fi - 1] > p[i - 1])
        wl = (double)(f[fi - 1] - p[i - 1]) / (double)(p[i] - p[i - 1]);
      if (fi + 1 < (int)f.size() && i + 1 < n && f[fi + 1] < p[i + 1])
        wr = (double)(p[i + 1] - f[fi + 1]) / (double)(p[i + 1] - p[i]);
      if (i == 0 && (int)p.size() != (int)f.size()) {
        wr *= 2.0;
      } else if (i == n - 1 && (int)p.size() != (int)f.size()) {
        wl *= 2.0;
      }
      t.wl = (float)wl;
      t.wr = (float)wr;
    }
    out.push_back(t);
  }
}

emcgpu_status build_multigrid(emcgpu_ctx *c, DeviceSim &d) {
  std::vector<std::vector<int>> PX, PY;
  std::vector<int> px(d.ext[0]), py(d.ext[1]);
  for (int i = 0; i < d.ext[0]; i++)
    px[i] = i;
  for (int j = 0; j < d.ext[1]; j++)
    py[j] = j;
  PX.push_back(px);
  PY.push_back(py);
  while ((int)PX.size() < kMaxLevels) {
    std::vector<int> qx = coarsen(PX.back()), qy = coarsen(PY.back());
    if (qx.size() == PX.back().size() && qy.size() == PY.back().size())
      break;
    PX.push_back(qx);
    PY.push_back(qy);
  }
  const int nl = (int)PX.size();
  d.lev.assign(nl, MGLevel{});
  d.tab.clear();
  d.lmask.clear();
  int off = 0;
  for (int l = 0; l < nl; l++) {
    MGLevel &L = d.lev[l];
    L.nx = (int)PX[l].size();
    L.ny = (int)PY[l].size();
    L.off = off;
    L.tx = (int)d.tab.size();
    dim_tables(PX[l], l + 1 < nl ? &PX[l + 1] : nullptr, l > 0 ? &PX[l - 1] : nullptr, d.tab);
    L.ty = (int)d.tab.size();
    dim_tables(PY[l], l + 1 < nl ? &PY[l + 1] : nullptr, l > 0 ? &PY[l - 1] : nullptr, d.tab);
    // Dirichlet nodes: coincident with an ohmic node of the device; a level
    // without any takes every node whose neighbourhood holds one
    std::vector<uint8_t> m(L.nx * L.ny, 0);
    int cnt = 0;
    for (int J = 0; J < L.ny; J++)
      for (int I = 0; I < L.nx; I++) {
        m[J * L.nx + I] = d.dmask[idx(d, PX[l][I], PY[l][J])];
        cnt += m[J * L.nx + I];
      }
    if (cnt == 0)
      for (int J = 0; J < L.ny; J++)
        for (int I = 0; I < L.nx; I++) {
          const int x0 = I > 0 ? PX[l][I - 1] + 1 : 0, x1 = I + 1 < L.nx ? PX[l][I + 1] - 1 : d.ext[0] - 1;
          const int y0 = J > 0 ? PY[l][J - 1] + 1 : 0, y1 = J + 1 < L.ny ? PY[l][J + 1] - 1 : d.ext[1] - 1;
          for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++)
              m[J * L.nx + I] |= d.dmask[idx(d, x, y)];
        }
    d.lmask.insert(d.lmask.end(), m.begin(), m.end());
    off += L.nx * L.ny;
  }
  d.mg_nodes = off;
  const int n0 = d.lev[0].nx * d.lev[0].ny, nc = off - n0;
  d.lds_res = off;
  d.lds_rc = d.lds_res + n0;
  d.lds_sc = d.lds_rc + nc;
  d.lds_tab = (int)(((size_t)(d.lds_sc + nc) * sizeof(float) + 15) & ~(size_t)15);
  d.lds_cmask = d.lds_tab + (int)(d.tab.size() * sizeof(DimTab));
  d.mg_lds = ((size_t)d.lds_cmask + nc + 15) & ~(size_t)15;
  // one word per thread for the fine-grid passes' off-grid slots
  d.lds_dummy = (int)(d.mg_lds / sizeof(float));
  d.mg_lds += kDevThreads * sizeof(float);
  d.wave_level = nl;
  for (int l = 1; l < nl; l++)
    if (d.lev[l].nx * d.lev[l].ny <= 256) {
      d.wave_level = l;
      break;
    }
  // exact inverse of the operator of the first coarse level with at most 256
  // nodes (rows / columns of Dirichlet nodes zero), row-major, FP32
  d.direct_level = 0;
  d.cinv_t.clear();
  // (a gate's Robin shift is not in this operator: with gates the coarse
  // levels are swept instead)
  if (d.wave_level < nl && !d.has_gate) {
    const int lc = d.wave_level;
    const MGLevel &C = d.lev[lc];
    const int n = C.nx * C.ny;
    const uint8_t *msk = d.lmask.data() + C.off;
    std::vector<int> act;
    std::vector<int> pos(n, -1);
    for (int k = 0; k < n; k++)
      if (!msk[k]) {
        pos[k] = (int)act.size();
        act.push_back(k);
      }
    const int m = (int)act.size();
    std::vector<double> A((size_t)m * m, 0.0), Inv((size_t)m * m, 0.0);
    for (int r = 0; r < m; r++) {
      const int k = act[r], j = k / C.nx, i = k - j * C.nx;
      const DimTab &tx = d.tab[C.tx + i], &ty = d.tab[C.ty + j];
      const double cl = d.hfx * ty.w * tx.inv_l, cr = d.hfx * ty.w * tx.inv_r;
      const double cd = d.hfy * tx.w * ty.inv_l, cu = d.hfy * tx.w * ty.inv_r;
      const int il = i > 0 ? k - 1 : k + 1, ir = i < C.nx - 1 ? k + 1 : k - 1;
      const int jd = j > 0 ? k - C.nx : k + C.nx, ju = j < C.ny - 1 ? k + C.nx : k - C.nx;
      A[(size_t)r * m + r] += cl + cr + cd + cu;
      const int nbk[4] = {il, ir, jd, ju};
      const double cf[4] = {cl, cr, cd, cu};
      for (int q = 0; q < 4; q++)
        if (pos[nbk[q]] >= 0)
          A[(size_t)r * m + pos[nbk[q]]] -= cf[q];
      Inv[(size_t)r * m + r] = 1.0;
    }
    // Gauss-Jordan elimination with partial pivoting
    bool ok = m > 0;
    for (int col = 0; col < m && ok; col++) {
      int piv = col;
      for (int r = col + 1; r < m; r++)
        if (std::fabs(A[(size_t)r * m + col]) > std::fabs(A[(size_t)piv * m + col]))
          piv = r;
      if (std::fabs(A[(size_t)piv * m + col]) < 1e-300) {
        ok = false;
        break;
      }
      if (piv != col)
        for (int q = 0; q < m; q++) {
          std::swap(A[(size_t)piv * m + q], A[(size_t)col * m + q]);
          std::swap(Inv[(size_t)piv * m + q], Inv[(size_t)col * m + q]);
        }
      const double inv = 1.0 / A[(size_t)col * m + col];
      for (int q = 0; q < m; q++) {
        A[(size_t)col * m + q] *= inv;
        Inv[(size_t)col * m + q] *= inv;
      }
      for (int r = 0; r < m; r++) {
        if (r == col)
          continue;
        const double f = A[(size_t)r * m + col];
        if (f == 0.0)
          continue;
        for (int q = 0; q < m; q++) {
          A[(size_t)r * m + q] -= f * A[(size_t)col * m + q];
          Inv[(size_t)r * m + q] -= f * Inv[(size_t)col * m + q];
        }
      }
    }
    if (ok) {
      d.cinv_t.assign((size_t)n * n, 0.0f);
      for (int r = 0; r < m; r++)
        for (int q = 0; q < m; q++)
          d.cinv_t[(size_t)act[r] * n + act[q]] = (float)Inv[(size_t)r * m + q];
      d.direct_level = lc;
    }
  }
  if (d.ext[0] > kDevThreads)
    return fail(c, EMCGPU_ERR_ARG,
                "a %d x %d grid exceeds the single-workgroup Poisson solver (at most %d nodes "
                "along x)",
                d.ext[0], d.ext[1], kDevThreads);
  auto clog2 = [](int v) {
    int l = 0;
    while ((1 << l) < v)
      ++l;
    return l;
  };
  d.hcw_log2 = clog2((d.ext[0] + 1) / 2);
  d.fcw_log2 = clog2(d.ext[0]);
  const int rh = kDevThreads >> d.hcw_log2, rf = kDevThreads >> d.fcw_log2;
  d.nph = (d.ext[1] + rh - 1) / rh;
  d.npf = (d.ext[1] + rf - 1) / rf;
  // levels beyond the LDS: the same layout in a global workspace (the
  // workgroup then reads them through L2)
  d.mg_global = d.mg_lds > (size_t)kDevMaxLds;
  // the Poisson passes' index maths (idiv) is exact below 2^22 nodes
  if ((size_t)d.mg_nodes + d.mg_lds / sizeof(float) >= ((size_t)1 << 22))
    return fail(c, EMCGPU_ERR_ARG, "a %d x %d grid exceeds the Poisson solver (%d nodes)",
                d.ext[0], d.ext[1], d.mg_nodes);
  return EMCGPU_OK;
}

// ---------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------
template <typename T>
emcgpu_status dalloc(emcgpu_ctx *c, T *&p, size_t n) {
  dfreet(p);
  HIPCHK(c, hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(n, 1)));
  return EMCGPU_OK;
}
emcgpu_status dallocv(emcgpu_ctx *c, void *&p, size_t bytes) {
  dfree(p);
  HIPCHK(c, hipMalloc(&p, std::max<size_t>(bytes, 16)));
  return EMCGPU_OK;
}
template <typename T>
emcgpu_status dput(emcgpu_ctx *c, T *&p, const std::vector<T> &v) {
  emcgpu_status st = dalloc(c, p, v.size());
  if (st)
    return st;
  if (!v.empty())
    HIPCHK(c, hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return EMCGPU_OK;
}

#define ST(x)                 \
  do {                        \
    emcgpu_status s_ = (x);   \
    if (s_ != EMCGPU_OK)      \
      return s_;              \
  } while (0)

emcgpu_status upload_device(emcgpu_ctx *c, DeviceSim &d) {
  const size_t nn = d.nn, rs = real_size(c);
  ST(dput(c, d.d_cnode, d.cnode));
  if (d.region_tables) {
    std::vector<int8_t> reg(nn);
    for (size_t k = 0; k < nn; k++)
      reg[k] = (int8_t)d.region[k];
    ST(dput(c, d.d_region, reg));
  }
  ST(dput(c, d.d_cslot, d.cslot));
  ST(dput(c, d.d_bcx, d.bcx));
  ST(dput(c, d.d_bcy, d.bcy));
  ST(dput(c, d.d_dmask, d.dmask));
  ST(dput(c, d.d_dopn, d.dopn));
  ST(dput(c, d.d_bc_eq, d.bc_eq));
  ST(dput(c, d.d_bc_neq, d.bc_neq));
  ST(dput(c, d.d_cfac, d.cfac));
  if (d.has_gate)
    ST(dput(c, d.d_gate, d.gate));
  ST(dput(c, d.d_slot_node, d.slot_node));
  ST(dput(c, d.d_slot_contact, d.slot_contact));
  ST(dput(c, d.d_slot_exp, d.slot_exp));
  ST(dput(c, d.d_tab, d.tab));
  ST(dput(c, d.d_lmask, d.lmask));
  const size_t ns = std::max<size_t>(d.slot_node.size(), 1);
  ST(dalloc(c, d.d_slot_kept, ns));
  ST(dalloc(c, d.d_slot_inj, ns));
  ST(dalloc(c, d.d_slot_total, ns));
  ST(dalloc(c, d.d_pot, nn));
  ST(dalloc(c, d.d_conc, nn));
  ST(dalloc(c, d.d_nrpart, nn));
  ST(dalloc(c, d.d_ex, nn));
  ST(dalloc(c, d.d_ey, nn));
  ST(dalloc(c, d.d_corr, nn));
  ST(dalloc(c, d.d_pot_t, nn));
  ST(dalloc(c, d.d_conc_t, nn));
  ST(dalloc(c, d.d_ex_t, 2 * nn));
  ST(dalloc(c, d.d_ey_t, nn));
  ST(dallocv(c, d.d_acc, 2 * nn * rs));
  d.n_pad = (int)((nn + kPad - 1) / kPad * kPad);
  const size_t np = d.n_pad;
  ST(dalloc(c, d.d_r64, np));
  ST(dalloc(c, d.d_x64, np));
  ST(dalloc(c, d.d_p64, np));
  ST(dalloc(c, d.d_q64, np));
  ST(dalloc(c, d.d_s64, np));
  ST(dalloc(c, d.d_rs0, np));
  if (d.mg_global)
    ST(dalloc(c, d.d_mg_ws, (d.mg_lds + sizeof(float) - 1) / sizeof(float)));
  {
    std::vector<float> w(np, 0.0f), hw(np, 0.0f);
    std::vector<float2> rs(np, make_float2(0.0f, 1e30f));
    for (size_t k = 0; k < nn; k++) {
      w[k] = d.dmask[k] ? 0.0f : 1.0f;
      const int ix = (int)(k % d.ext[0]), iy = (int)(k / d.ext[0]);
      hw[k] = (ix == 0 || ix == d.ext[0] - 1 ? 0.5f : 1.0f) *
              (iy == 0 || iy == d.ext[1] - 1 ? 0.5f : 1.0f);
    }
    ST(dput(c, d.d_w0, w));
    ST(dput(c, d.d_hw0, hw));
    ST(dput(c, d.d_cinv_t, d.cinv_t));
    HIPCHK(c, hipMemcpy(d.d_rs0, rs.data(), np * sizeof(float2), hipMemcpyHostToDevice));
    std::vector<double> z(np, 0.0);
    for (double *v : {d.d_r64, d.d_x64, d.d_p64, d.d_q64, d.d_s64})
      HIPCHK(c, hipMemcpy(v, z.data(), np * sizeof(double), hipMemcpyHostToDevice));
  }
  ST(dalloc(c, d.d_state, 1));
  ST(dallocv(c, d.d_fargs, 4 * std::max(sizeof(DevFieldArgs<float>), sizeof(DevFieldArgs<double>))));
  ST(dalloc(c, d.d_iters, 1));
  ST(dalloc(c, d.d_resid, 1));
  HIPCHK(c, hipMemset(d.d_pot, 0, sizeof(double) * nn));
  HIPCHK(c, hipMemset(d.d_ex, 0, sizeof(double) * nn));
  HIPCHK(c, hipMemset(d.d_ey, 0, sizeof(double) * nn));
  HIPCHK(c, hipMemset(d.d_corr, 0, sizeof(double) * nn));
  HIPCHK(c, hipMemset(d.d_acc, 0, 2 * nn * rs));
  DevState s{};
  HIPCHK(c, hipMemcpy(d.d_state, &s, sizeof s, hipMemcpyHostToDevice));
  // step kernel: a block-private deposit grid when it fits the LDS (the
  // rank kernel holds the per-wave slot counts, at most 32 KB)
  const size_t grid_bytes = ((nn * rs) + 15) & ~(size_t)15;
  d.lds_dep = grid_bytes <= (size_t)kDevMaxLds;
  d.step_lds = d.lds_dep ? grid_bytes : 0;
  hipDeviceProp_t prop;
  HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
  // one step block per CU (its waves share the CU's chunks)
  const int per_cu = c->precision == EMCGPU_FP32 ? dev_step_blocks_per_cu<float>(d.step_lds)
                                                 : dev_step_blocks_per_cu<double>(d.step_lds);
  if (per_cu < 1)
    return fail(c, EMCGPU_ERR_ARG, "the step kernel does not fit a CU");
  d.blocks = prop.multiProcessorCount;
  d.n_seg = prop.multiProcessorCount * kDevSegPerCU;
  const size_t dep_rows = d.lds_dep ? (size_t)d.blocks : 1;
  ST(dallocv(c, d.d_dep, dep_rows * nn * rs));
  d.rows_per_chunk = 16;
  d.n_chunks = (int)((dep_rows + d.rows_per_chunk - 1) / d.rows_per_chunk);
  ST(dalloc(c, d.d_dpart, (size_t)d.n_chunks * nn));
  HIPCHK(c, hipMemset(d.d_dep, 0, dep_rows * nn * rs));
  ST(dalloc(c, d.d_seg_cnt, ns * d.n_seg));
  ST(dalloc(c, d.d_seg_off, ns * d.n_seg));
  ST(dalloc(c, d.d_seg_nst, d.n_seg));
  ST(dalloc(c, d.d_chan_part, (size_t)d.blocks));
  ST(dalloc(c, d.d_exit_part, (size_t)d.blocks * std::max(d.nContacts, 1)));
  ST(dalloc(c, d.d_chan_corr, 1));
  ST(dalloc(c, d.d_avg_pot, nn));
  ST(dalloc(c, d.d_avg_conc, nn));
  HIPCHK(c, hipMemset(d.d_chan_corr, 0, sizeof(double)));
  HIPCHK(c, hipMemset(d.d_avg_pot, 0, sizeof(double) * nn));
  HIPCHK(c, hipMemset(d.d_avg_conc, 0, sizeof(double) * nn));
  d.on_gpu = true;
  return EMCGPU_OK;
}

emcgpu_status ensure_hist(emcgpu_ctx *c, DeviceSim &d, int rows) {
  if (rows > d.hist_rows) {
    const size_t nc = std::max(d.nContacts, 1);
    ST(dalloc(c, d.d_hexits, rows * nc));
    ST(dalloc(c, d.d_hdels, rows * nc));
    ST(dalloc(c, d.d_hinjs, rows * nc));
    ST(dalloc(c, d.d_hiters, rows));
    ST(dalloc(c, d.d_hnpart, rows));
    ST(dalloc(c, d.d_hresid, rows));
    ST(dalloc(c, d.d_hcheck, rows));
    ST(dalloc(c, d.d_hchan, rows));
    d.hist_rows = rows;
    if (d.gexec) {
      (void)hipGraphExecDestroy(d.gexec);
      d.gexec = nullptr;
    }
  }
  return EMCGPU_OK;
}

DevHist hist_of(const DeviceSim &d) {
  DevHist h{};
  h.exits = d.d_hexits;
  h.dels = d.d_hdels;
  h.injs = d.d_hinjs;
  h.n_part = d.d_hnpart;
  h.iters = d.d_hiters;
  h.resid = d.d_hresid;
  h.checksum = d.d_hcheck;
  h.chan = d.d_hchan;
  h.rows = d.hist_rows;
  return h;
}

template <typename R>
DevParticles<R> parts(const DeviceSim &d, int b) {
  DevParticles<R> p;
  p.x = (R *)d.px[b];
  p.y = (R *)d.py[b];
  p.kx = (R *)d.pkx[b];
  p.ky = (R *)d.pky[b];
  p.kz = (R *)d.pkz[b];
  p.E = (R *)d.pE[b];
  p.tau = (R *)d.ptau[b];
  p.vs = d.pvs[b];
  p.pid = d.ppid[b];
  return p;
}

emcgpu_status read_state(emcgpu_ctx *c, const DeviceSim &d, DevState &s) {
  HIPCHK(c, hipMemcpyAsync(&s, d.d_state, sizeof s, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EMCGPU_OK;
}
emcgpu_status write_state(emcgpu_ctx *c, const DeviceSim &d, const DevState &s) {
  HIPCHK(c, hipMemcpyAsync(d.d_state, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EMCGPU_OK;
}

// grows both particle buffer sets to hold need entries (current contents kept)
emcgpu_status ensure_capacity(emcgpu_ctx *c, DeviceSim &d, long long need, long long keep) {
  if (need <= d.capacity)
    return EMCGPU_OK;
  const long long cap = std::max<long long>(need + need / 4, 4096);
  const size_t rs = real_size(c);
  void **arrs[7] = {d.px, d.py, d.pkx, d.pky, d.pkz, d.pE, d.ptau};
  for (int b = 0; b < 2; b++) {
    for (void **a : arrs) {
      void *n = nullptr;
      HIPCHK(c, hipMalloc(&n, cap * rs));
      if (a[b] && keep > 0 && b == d.cur)
        HIPCHK(c, hipMemcpy(n, a[b], keep * rs, hipMemcpyDeviceToDevice));
      dfree(a[b]);
      a[b] = n;
    }
    uint32_t *nv = nullptr, *np = nullptr;
    HIPCHK(c, hipMalloc((void **)&nv, cap * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc((void **)&np, cap * sizeof(uint32_t)));
    if (d.pvs[b] && keep > 0 && b == d.cur) {
      HIPCHK(c, hipMemcpy(nv, d.pvs[b], keep * sizeof(uint32_t), hipMemcpyDeviceToDevice));
      HIPCHK(c, hipMemcpy(np, d.ppid[b], keep * sizeof(uint32_t), hipMemcpyDeviceToDevice));
    }
    dfreet(d.pvs[b]);
    dfreet(d.ppid[b]);
    d.pvs[b] = nv;
    d.ppid[b] = np;
  }
  ST(dalloc(c, d.d_stage, cap));
  ST(dalloc(c, d.d_pslot, cap));
  ST(dalloc(c, d.d_blk, cap / 1024 + 2));
  d.capacity = cap;
  DevState s{};
  ST(read_state(c, d, s));
  s.capacity = cap;
  ST(write_state(c, d, s));
  if (d.gexec) {
    (void)hipGraphExecDestroy(d.gexec);
    d.gexec = nullptr;
  }
  return EMCGPU_OK;
}

template <typename R>
DevGeom<R> geom_nm(const DeviceSim &d) {
  DevGeom<R> g;
  g.nx = d.ext[0];
  g.ny = d.ext[1];
  g.nn = d.nn;
  g.pm = d.pm;
  g.hx = (R)(d.h[0] * kNm);
  g.hy = (R)(d.h[1] * kNm);
  g.maxx = (R)(d.maxPos[0] * kNm);
  g.maxy = (R)(d.maxPos[1] * kNm);
  g.cnode = d.d_cnode;
  g.cslot = d.d_cslot;
  g.acc = (const R *)d.d_acc;
  g.region = d.region_tables ? d.d_region : nullptr;
  return g;
}

// geometry in metres for the known-answer entry points
DevGeom<double> geom_m(const DeviceSim &d, const double *acc) {
  DevGeom<double> g;
  g.nx = d.ext[0];
  g.ny = d.ext[1];
  g.nn = d.nn;
  g.pm = d.pm;
  g.hx = d.h[0];
  g.hy = d.h[1];
  g.maxx = d.maxPos[0];
  g.maxy = d.maxPos[1];
  g.cnode = d.d_cnode;
  g.cslot = d.d_cslot;
  g.acc = acc;
  return g;
}

// force / hbar scale of the acceleration grid: charge -q, k in 1/nm, t in fs
constexpr double kAccScale = -phys::q / phys::hbar * 1e-9 * 1e-15;

template <typename R>
DevFieldArgs<R> field_args(const emcgpu_ctx *c, const DeviceSim &d) {
  (void)c;
  DevFieldArgs<R> a;
  std::memset(&a, 0, sizeof a);
  a.nx = d.ext[0];
  a.ny = d.ext[1];
  a.nn = d.nn;
  a.pm = d.pm;
  a.hfx = d.hfx;
  a.hfy = d.hfy;
  a.hp = d.hp;
  a.vt = d.Vt;
  a.hxm = d.h[0];
  a.hym = d.h[1];
  a.tol = d.tol;
  a.max_iter = d.max_iter;
  a.pot = d.d_pot;
  a.dopn = d.d_dopn;
  a.conc = d.d_conc;
  a.bc_eq = d.d_bc_eq;
  a.bc_neq = d.d_bc_neq;
  a.gate = d.d_gate;
  a.dmask = d.d_dmask;
  a.bcx = d.d_bcx;
  a.bcy = d.d_bcy;
  a.ex = d.d_ex;
  a.ey = d.d_ey;
  a.acc = (R *)d.d_acc;
  a.acc_scale = kAccScale;
  a.r64 = d.d_r64;
  a.x64 = d.d_x64;
  a.p64 = d.d_p64;
  a.q64 = d.d_q64;
  a.s64 = d.d_s64;
  a.rs0 = d.d_rs0;
  a.w0 = d.d_w0;
  a.hw0 = d.d_hw0;
  a.cinv = d.d_cinv_t;
  a.direct_level = d.direct_level;
  a.n_pad = d.n_pad;
  a.max_vcycles = 400;
  a.max_pcg = 60;
  a.n_levels = (int)d.lev.size();
  for (size_t l = 0; l < d.lev.size(); l++)
    a.lev[l] = d.lev[l];
  a.tab = d.d_tab;
  a.n_tab = (int)d.tab.size();
  a.lmask = d.d_lmask;
  a.n_coarse = d.mg_nodes - d.nn;
  a.wave_level = d.wave_level;
  a.nu_pre = 2;
  a.nu_post = 2;
  a.nu_coarse = 4;
  a.lds_res = d.lds_res;
  a.lds_rc = d.lds_rc;
  a.lds_sc = d.lds_sc;
  a.lds_tab = d.lds_tab;
  a.lds_cmask = d.lds_cmask;
  a.lds_dummy = d.lds_dummy;
  a.mg_ws = d.mg_global ? d.d_mg_ws : nullptr;
  a.hcw_log2 = d.hcw_log2;
  a.nph = d.nph;
  a.fcw_log2 = d.fcw_log2;
  a.npf = d.npf;
  a.state = d.d_state;
  a.hist = hist_of(d);
  return a;
}

template <typename R>
DevStepArgs<R> step_args(const emcgpu_ctx *c, const DeviceSim &d) {
  DevStepArgs<R> a;
  std::memset(&a, 0, sizeof a);
  a.p = parts<R>(d, d.cur);
  a.g = geom_nm<R>(d);
  a.inst = (const InstDev<R> *)device_params(c);
  a.cum = (const R *)device_tables(c);
  a.state = d.d_state;
  a.hist = hist_of(d);
  a.key0 = (uint32_t)d.seed;
  a.key1 = (uint32_t)(d.seed >> 32);
  a.n_contacts = d.nContacts;
  a.n_slots = (int)d.slot_node.size();
  a.dep = (R *)d.d_dep;
  a.seg_cnt = d.d_seg_cnt;
  a.seg_nst = d.d_seg_nst;
  a.stage = d.d_stage;
  a.pslot = d.d_pslot;
  a.exit_part = d.d_exit_part;
  a.lds_dep = d.lds_dep ? 1 : 0;
  a.move = 1;
  a.chan = d.chan;
  a.chan_x0 = (R)d.chan_x0;
  a.chan_x1 = (R)d.chan_x1;
  a.chan_part = d.d_chan_part;
  return a;
}

template <typename R>
DevContactArgs<R> contact_args(const emcgpu_ctx *c, const DeviceSim &d) {
  DevContactArgs<R> a;
  std::memset(&a, 0, sizeof a);
  a.p = parts<R>(d, d.cur);
  a.g = geom_nm<R>(d);
  a.inst = (const InstDev<R> *)device_params(c);
  a.state = d.d_state;
  a.hist = hist_of(d);
  a.key0 = (uint32_t)d.seed;
  a.key1 = (uint32_t)(d.seed >> 32);
  a.n_contacts = d.nContacts;
  a.n_slots = (int)d.slot_node.size();
  a.n_segments = d.n_seg;
  a.n_blocks_dep = d.lds_dep ? d.blocks : 1;
  a.seg_cnt = d.d_seg_cnt;
  a.seg_off = d.d_seg_off;
  a.seg_nst = d.d_seg_nst;
  a.stage = d.d_stage;
  a.slot_node = d.d_slot_node;
  a.slot_contact = d.d_slot_contact;
  a.slot_exp = d.d_slot_exp;
  a.slot_kept = d.d_slot_kept;
  a.slot_inj = d.d_slot_inj;
  a.slot_total = d.d_slot_total;
  a.exit_part = d.d_exit_part;
  a.n_step_blocks = d.blocks;
  a.corr = d.d_corr;
  a.dep = (R *)d.d_dep;
  a.dpart = d.d_dpart;
  a.rows_per_chunk = d.rows_per_chunk;
  a.n_chunks = d.n_chunks;
  a.nrpart = d.d_nrpart;
  a.conc = d.d_conc;
  a.cfac = d.d_cfac;
  a.chan = d.chan;
  a.chan_x0 = (R)d.chan_x0;
  a.chan_x1 = (R)d.chan_x1;
  a.chan_part = d.d_chan_part;
  a.n_chan_blocks = d.blocks;
  a.chan_corr = d.d_chan_corr;
  a.chan_scale = d.chan_scale;
  a.accumulate = d.accumulate;
  a.pot = d.d_pot;
  a.avg_pot = d.d_avg_pot;
  a.avg_conc = d.d_avg_conc;
  return a;
}

emcgpu_status need_device(emcgpu_ctx *c, bool initialized = false) {
  if (!c)
    return EMCGPU_ERR_ARG;
  if (!c->dsim)
    return fail(c, EMCGPU_ERR_STATE, "no device (call emcgpu_device_setup first)");
  if (!c->dsim->on_gpu)
    return fail(c, EMCGPU_ERR_STATE, "the device path needs a GPU context");
  if (initialized && !c->dsim->initialized)
    return fail(c, EMCGPU_ERR_STATE, "no device ensemble (call emcgpu_device_init first)");
  return EMCGPU_OK;
}

emcgpu_status need_model(emcgpu_ctx *c, double dt_s) {
  if (!c->inst[0].has_tables)
    return fail(c, EMCGPU_ERR_STATE, "instance 0 has no scatter tables");
  for (const auto &m : c->inst[0].mechs)
    if (m.kind == EMCGPU_MECH_HOT_FROEHLICH_AB || m.kind == EMCGPU_MECH_HOT_FROEHLICH_EM ||
        m.kind == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB ||
        m.kind == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_EM)
      return fail(c, EMCGPU_ERR_ARG, "hot-phonon mechanisms are not supported in device runs");
  if (c->inst[0].plasma())
    return fail(c, EMCGPU_ERR_ARG, "plasma screening is not supported in device runs");
  return upload_model(c, dt_s);
}

// per-region tables: every region's instance has tables, the valleys of
// instance 0 and no mechanism the device path refuses
emcgpu_status need_regions(emcgpu_ctx *c, const DeviceSim &d) {
  if (!d.region_tables)
    return EMCGPU_OK;
  for (int r = 0; r < d.n_regions; r++) {
    const auto &in = c->inst[r];
    if (!in.has_tables)
      return fail(c, EMCGPU_ERR_STATE, "instance %d (tables of region %d) has no tables", r, r);
    bool same = in.valleys.size() == c->inst[0].valleys.size();
    for (size_t v = 0; same && v < in.valleys.size(); v++) {
      const emcgpu_valley_desc &a = in.valleys[v], &b = c->inst[0].valleys[v];
      same = a.kind == b.kind && a.particle_mass == b.particle_mass &&
             a.degeneracy == b.degeneracy && a.alpha == b.alpha &&
             a.bottom_energy == b.bottom_energy &&
             std::memcmp(a.rel_mass, b.rel_mass, sizeof a.rel_mass) == 0 &&
             std::memcmp(a.axes, b.axes, sizeof a.axes) == 0;
    }
    if (!same)
      return fail(c, EMCGPU_ERR_ARG, "instance %d: per-region tables need the valleys of instance 0",
                  r);
    for (const auto &m : in.mechs)
      if (m.kind == EMCGPU_MECH_HOT_FROEHLICH_AB || m.kind == EMCGPU_MECH_HOT_FROEHLICH_EM ||
          m.kind == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB ||
          m.kind == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_EM)
        return fail(c, EMCGPU_ERR_ARG, "hot-phonon mechanisms are not supported in device runs");
    if (in.plasma())
      return fail(c, EMCGPU_ERR_ARG, "plasma screening is not supported in device runs");
  }
  return EMCGPU_OK;
}

// device slot of the field-kernel arguments: 0 steps, 1 the first step
// (contacts reset), 2 the other calls
template <typename R>
DevFieldArgs<R> *fargs_slot(const DeviceSim &d, int slot) {
  return (DevFieldArgs<R> *)d.d_fargs + slot;
}

// writes the arguments of a field-kernel slot (synchronous: never inside a
// captured graph, whose launches only reference the slot)
template <typename R>
emcgpu_status put_fargs(emcgpu_ctx *c, DeviceSim &d, int slot, const DevFieldArgs<R> &a) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(fargs_slot<R>(d, slot), &a, sizeof(a), hipMemcpyHostToDevice));
  return EMCGPU_OK;
}

// one Poisson / E-field launch on the given potential and outputs
template <typename R>
emcgpu_status run_field(emcgpu_ctx *c, DeviceSim &d, DevFieldArgs<R> a) {
  ST(put_fargs<R>(c, d, 2, a));
  HIPCHK(c, dev_launch_field<R>(fargs_slot<R>(d, 2), d.mg_lds, d.mg_global, c->stream));
  return EMCGPU_OK;
}

// dense corrections per step solve (each one shrinks the residual by the
// FP32 rounding of the inverse, ~1e-5 or better; the multigrid solve after
// them checks the stopping rule)
constexpr int kGreenIters = 2;
// the hole term h^2 p the inverse leaves out, relative to the diagonal
constexpr double kGreenShiftRel = 1e-9;

// G = (H A0)^-1 of the active nodes for the dense Newton corrections: the
// operator of the Newton step without the hole term, weighted by the
// control volumes (symmetric positive definite), assembled on the host,
// inverted in FP64 with rocSOLVER (Cholesky, potrf + potri), stored FP32
// as [nn][ld] with zero rows and columns at the Dirichlet nodes.  Any
// failure leaves the multigrid path alone in charge (green_state -1).
emcgpu_status green_checks_alloc(emcgpu_ctx *c, DeviceSim &d);

emcgpu_status green_build(emcgpu_ctx *c, DeviceSim &d) {
  if (d.green_state != 0)
    return EMCGPU_OK;
  d.green_state = -1;
  const int nx = d.ext[0], ny = d.ext[1], nn = (int)d.nn;
  std::vector<int> idx(nn, -1);
  int na = 0;
  for (int k = 0; k < nn; k++)
    if (!d.dmask[k])
      idx[k] = na++;
  const int ld = (nn + 255) / 256 * 256;
  if (na == 0 || na > kGreenMaxActive || (size_t)ld * sizeof(float) > 160 * 1024)
    return EMCGPU_OK;
  // entries of H A0, row by row (the stencil of k_dev_field: mirrored
  // neighbours at the boundaries, Dirichlet neighbours dropped)
  std::map<std::pair<int, int>, double> ent;
  for (int k = 0; k < nn; k++) {
    if (idx[k] < 0)
      continue;
    const int i = k % nx, j = k / nx;
    const double hw = (i == 0 || i == nx - 1 ? 0.5 : 1.0) * (j == 0 || j == ny - 1 ? 0.5 : 1.0);
    ent[{idx[k], idx[k]}] += hw * 2.0 * (d.hfx + d.hfy);
    if (d.has_gate)  // the gate's Robin shift
      ent[{idx[k], idx[k]}] += hw * d.gate[3 * (size_t)k];
    const int nb[4] = {i > 0 ? k - 1 : k + 1, i < nx - 1 ? k + 1 : k - 1,
                       j > 0 ? k - nx : k + nx, j < ny - 1 ? k + nx : k - nx};
    for (int q = 0; q < 4; q++)
      if (idx[nb[q]] >= 0)
        ent[{idx[k], idx[nb[q]]}] -= hw * (q < 2 ? d.hfx : d.hfy);
  }
  std::vector<int> rows, cols;
  std::vector<double> vals;
  for (const auto &e : ent) {
    const auto t = ent.find({e.first.second, e.first.first});
    if (t == ent.end() || std::fabs(t->second - e.second) > 1e-12 * std::fabs(e.second))
      return EMCGPU_OK;  // not symmetric: no Cholesky
    rows.push_back(e.first.first);
    cols.push_back(e.first.second);
    vals.push_back(e.second);
  }
  double *dA = nullptr, *dvals = nullptr;
  int *drows = nullptr, *dcols = nullptr, *didx = nullptr;
  rocblas_int *dinfo = nullptr;
  float *dG = nullptr;
  rocblas_handle h = nullptr;
  auto release = [&]() {
    dfreet(dA);
    dfreet(dvals);
    dfreet(drows);
    dfreet(dcols);
    dfreet(didx);
    dfreet(dinfo);
    dfreet(dG);
    if (h)
      (void)rocblas_destroy_handle(h);
    (void)hipGetLastError();  // an allocation failure here only disables the path
  };
  const size_t n2 = (size_t)na * na;
  bool ok = hipMalloc((void **)&dA, n2 * sizeof(double)) == hipSuccess &&
            hipMalloc((void **)&dvals, vals.size() * sizeof(double)) == hipSuccess &&
            hipMalloc((void **)&drows, rows.size() * sizeof(int)) == hipSuccess &&
            hipMalloc((void **)&dcols, cols.size() * sizeof(int)) == hipSuccess &&
            hipMalloc((void **)&didx, (size_t)nn * sizeof(int)) == hipSuccess &&
            hipMalloc((void **)&dinfo, sizeof(rocblas_int)) == hipSuccess &&
            hipMalloc((void **)&dG, (size_t)nn * ld * sizeof(float)) == hipSuccess;
  if (!ok) {
    release();
    return EMCGPU_OK;
  }
  HIPCHK(c, hipMemsetAsync(dA, 0, n2 * sizeof(double), c->stream));
  HIPCHK(c, hipMemcpy(dvals, vals.data(), vals.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(drows, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(dcols, cols.data(), cols.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(didx, idx.data(), (size_t)nn * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(c, dev_launch_green_scatter(drows, dcols, dvals, (int)vals.size(), dA, na, c->stream));
  rocblas_int info = 0;
  ok = rocblas_create_handle(&h) == rocblas_status_success &&
       rocblas_set_stream(h, c->stream) == rocblas_status_success &&
       rocsolver_dpotrf(h, rocblas_fill_lower, na, dA, na, dinfo) == rocblas_status_success &&
       hipStreamSynchronize(c->stream) == hipSuccess &&
       hipMemcpy(&info, dinfo, sizeof info, hipMemcpyDeviceToHost) == hipSuccess && info == 0 &&
       rocsolver_dpotri(h, rocblas_fill_lower, na, dA, na, dinfo) == rocblas_status_success &&
       hipStreamSynchronize(c->stream) == hipSuccess &&
       hipMemcpy(&info, dinfo, sizeof info, hipMemcpyDeviceToHost) == hipSuccess && info == 0;
  if (!ok) {
    release();
    return EMCGPU_OK;
  }
  HIPCHK(c, dev_launch_green_expand(dA, na, didx, nn, ld, dG, c->stream));
  HIPCHK(c, dev_green_lds((size_t)ld * sizeof(float)));
  ST(green_checks_alloc(c, d));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d.d_green = dG;
  dG = nullptr;
  release();
  int cus = 0;
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  const int rows_per_block = kGreenThreads / 64;
  d.green_blocks = std::max(1, std::min(2 * cus, (nn + rows_per_block - 1) / rows_per_block));
  d.green_ld = ld;
  d.green_state = 1;
  return EMCGPU_OK;
}

GreenArgs green_args(const DeviceSim &d) {
  GreenArgs g;
  std::memset(&g, 0, sizeof g);
  g.nx = d.ext[0];
  g.ny = d.ext[1];
  g.nn = (int)d.nn;
  g.ld = d.green_ld;
  g.hfx = d.hfx;
  g.hfy = d.hfy;
  g.hp = d.hp;
  g.pot = d.d_pot;
  g.conc = d.d_conc;
  g.dopn = d.d_dopn;
  g.gate = d.d_gate;
  g.w0 = d.d_w0;
  g.hw0 = d.d_hw0;
  g.green = d.d_green;
  g.g = d.d_r64;  // scratch: the field kernel after the corrections rewrites it
  g.flag = d.d_green_flag;
  g.shift_max = kGreenShiftRel * 2.0 * (d.hfx + d.hfy);
  g.acc = d.d_green_acc;
  g.tol = d.tol;
  g.state = d.d_state;
  g.interval = d.interval;
  g.use_fd = d.fd_use;
  if (d.fd_use)
    g.fd = d.fd;
  return g;
}

// the residual checks of the corrections (shared by both inverses)
emcgpu_status green_checks_alloc(emcgpu_ctx *c, DeviceSim &d) {
  if (!d.d_green_flag) {
    HIPCHK(c, hipMalloc((void **)&d.d_green_flag, sizeof(unsigned long long)));
    HIPCHK(c, hipMemset(d.d_green_flag, 0, sizeof(unsigned long long)));
  }
  if (!d.d_green_acc) {
    const size_t acc_bytes = 2 * (kGreenIters + 1) * sizeof(unsigned long long);
    HIPCHK(c, hipMalloc((void **)&d.d_green_acc, acc_bytes));
    HIPCHK(c, hipMemset(d.d_green_acc, 0, acc_bytes));
  }
  return EMCGPU_OK;
}

// The separable inverse.  The Newton operator without the hole term,
// A0 = -J0, acts per dimension: hF_d (2 x_k - x_prev - x_next) with the
// mirrored neighbour at a Neumann boundary and Dirichlet neighbours dropped,
// plus the gates' shifts.  When the Dirichlet nodes are exactly the union of
// whole columns and whole rows, and the shift of the x-boundary gates
// depends only on the row and that of the y-boundary gates only on the
// column, A0 is the Kronecker sum Ax (+) Ay of two 1-D operators over the
// active columns and rows.  Each 1-D operator is symmetric under the
// control-volume weights w (1/2 at a boundary node): S = W^1/2 A W^-1/2 =
// Q L Q^T (rocSOLVER dsyevd, FP64), so A = B L P with B = W^-1/2 Q and
// P = Q^T W^1/2 = B^-1.  Not separable, not positive definite, or any
// failure: fd_state -1 (the dense inverse or multigrid then takes over).
emcgpu_status fd_build(emcgpu_ctx *c, DeviceSim &d) {
  if (d.fd_state != 0)
    return EMCGPU_OK;
  d.fd_state = -1;
  const int nx = d.ext[0], ny = d.ext[1];
  std::vector<uint8_t> colD(nx, 1), rowD(ny, 1);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      if (!d.dmask[idx(d, i, j)]) {
        colD[i] = 0;
        rowD[j] = 0;
      }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      if ((d.dmask[idx(d, i, j)] != 0) != (colD[i] || rowD[j]))
        return EMCGPU_OK;  // a Dirichlet node set that is no union of lines
  std::vector<int> ix, iy;
  for (int i = 0; i < nx; i++)
    if (!colD[i])
      ix.push_back(i);
  for (int j = 0; j < ny; j++)
    if (!rowD[j])
      iy.push_back(j);
  const int mx = (int)ix.size(), my = (int)iy.size();
  if (mx == 0 || my == 0)
    return EMCGPU_OK;
  // gate shifts: the x part must not vary along y, the y part not along x
  std::vector<double> gx(nx, 0.0), gy(ny, 0.0);
  if (d.has_gate) {
    for (int i : ix)
      gx[i] = d.gate_xy[2 * (size_t)idx(d, i, iy[0])];
    for (int j : iy)
      gy[j] = d.gate_xy[2 * (size_t)idx(d, ix[0], j) + 1];
    for (int j : iy)
      for (int i : ix) {
        const size_t k = (size_t)idx(d, i, j);
        if (d.gate_xy[2 * k] != gx[i] || d.gate_xy[2 * k + 1] != gy[j])
          return EMCGPU_OK;
      }
  }
  // S of one dimension (n nodes, active list act, Dirichlet lines dl, h-factor
  // hf, shifts g), column-major, with the weights' square roots in sw
  auto build_1d = [](int n, const std::vector<int> &act, const std::vector<uint8_t> &dl,
                     double hf, const std::vector<double> &g, std::vector<double> &S,
                     std::vector<double> &sw) {
    const int m = (int)act.size();
    std::vector<int> pos(n, -1);
    for (int a = 0; a < m; a++)
      pos[act[a]] = a;
    std::vector<double> A((size_t)m * m, 0.0);
    sw.assign(m, 1.0);
    for (int a = 0; a < m; a++) {
      const int i = act[a];
      A[(size_t)a * m + a] += 2.0 * hf + g[i];
      const int nb[2] = {i > 0 ? i - 1 : i + 1, i < n - 1 ? i + 1 : i - 1};
      for (int q = 0; q < 2; q++)
        if (!dl[nb[q]])
          A[(size_t)a * m + pos[nb[q]]] -= hf;
      sw[a] = (i == 0 || i == n - 1) ? std::sqrt(0.5) : 1.0;
    }
    S.assign((size_t)m * m, 0.0);
    for (int a = 0; a < m; a++)
      for (int b = 0; b < m; b++)  // row a, column b: stored at [b * m + a]
        S[(size_t)b * m + a] = sw[a] * A[(size_t)a * m + b] / sw[b];
    for (int a = 0; a < m; a++)  // symmetric up to rounding: take the mean
      for (int b = a + 1; b < m; b++) {
        const double v = 0.5 * (S[(size_t)b * m + a] + S[(size_t)a * m + b]);
        S[(size_t)b * m + a] = S[(size_t)a * m + b] = v;
      }
  };
  std::vector<double> Sx, Sy, swx, swy;
  build_1d(nx, ix, colD, d.hfx, gx, Sx, swx);
  build_1d(ny, iy, rowD, d.hfy, gy, Sy, swy);
  const std::vector<double> Sx0 = Sx, Sy0 = Sy;
  // the eigenpairs must solve S q = l q (S is tridiagonal: O(m^2) check)
  auto eig_ok = [](const std::vector<double> &S, const std::vector<double> &Q,
                   const std::vector<double> &lam, int m) {
    double smax = 0.0;
    for (double v : lam)
      smax = std::max(smax, std::fabs(v));
    for (int k = 0; k < m; k++)
      for (int a = 0; a < m; a++) {
        double r = -lam[k] * Q[(size_t)k * m + a];
        for (int b = std::max(0, a - 1); b <= std::min(m - 1, a + 1); b++)
          r += S[(size_t)b * m + a] * Q[(size_t)k * m + b];
        if (!(std::fabs(r) <= 1e-10 * smax))
          return false;
      }
    return true;
  };
  // eigen-decompositions on the device
  std::vector<double> lx(mx), ly(my);
  {
    double *dA = nullptr, *dD = nullptr, *dE = nullptr;
    rocblas_int *dinfo = nullptr;
    rocblas_handle h = nullptr;
    const int mm = std::max(mx, my);
    bool ok = hipMalloc((void **)&dA, (size_t)mm * mm * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&dD, (size_t)mm * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&dE, (size_t)mm * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&dinfo, sizeof(rocblas_int)) == hipSuccess &&
              rocblas_create_handle(&h) == rocblas_status_success &&
              rocblas_set_stream(h, c->stream) == rocblas_status_success;
    auto eig = [&](std::vector<double> &S, int m, std::vector<double> &lam) {
      rocblas_int info = 1;
      return hipMemcpy(dA, S.data(), (size_t)m * m * sizeof(double), hipMemcpyHostToDevice) ==
                 hipSuccess &&
             rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, m, dA, m, dD, dE,
                              dinfo) == rocblas_status_success &&
             // the context's stream does not order against the null-stream copies
             hipStreamSynchronize(c->stream) == hipSuccess &&
             hipMemcpy(&info, dinfo, sizeof info, hipMemcpyDeviceToHost) == hipSuccess &&
             info == 0 &&
             hipMemcpy(S.data(), dA, (size_t)m * m * sizeof(double), hipMemcpyDeviceToHost) ==
                 hipSuccess &&
             hipMemcpy(lam.data(), dD, (size_t)m * sizeof(double), hipMemcpyDeviceToHost) ==
                 hipSuccess;
    };
    ok = ok && eig(Sx, mx, lx) && eig(Sy, my, ly);
    dfreet(dA);
    dfreet(dD);
    dfreet(dE);
    dfreet(dinfo);
    if (h)
      (void)rocblas_destroy_handle(h);
    (void)hipGetLastError();  // a failure here only disables the path
    if (!ok)
      return EMCGPU_OK;
  }
  if (!eig_ok(Sx0, Sx, lx, mx) || !eig_ok(Sy0, Sy, ly, my))
    return EMCGPU_OK;
  // positive definite (the smallest eigenvalue of A0 is lx[0] + ly[0])
  if (!(lx[0] + ly[0] > 1e-12 * (lx[mx - 1] + ly[my - 1])))
    return EMCGPU_OK;
  // Q (column-major: Q[a][k] at [k * m + a]) -> the kernels' matrices
  const size_t nxx = (size_t)mx * mx, nyy = (size_t)my * my, nyx = (size_t)my * mx;
  std::vector<double> mats(2 * nyy + 2 * nxx + 2 * nyx, 0.0);
  double *py = mats.data(), *by = py + nyy, *pxt = by + nyy, *bxt = pxt + nxx,
         *inv_den = bxt + nxx;
  for (int j = 0; j < my; j++)
    for (int b = 0; b < my; b++) {
      const double q = Sy[(size_t)j * my + b];  // Q[b][j]
      py[(size_t)j * my + b] = q * swy[b];      // Py[j][b] = Q[b][j] sqrt(w_b)
      by[(size_t)b * my + j] = q / swy[b];      // By[b][j] = Q[b][j] / sqrt(w_b)
    }
  for (int k = 0; k < mx; k++)
    for (int a = 0; a < mx; a++) {
      const double q = Sx[(size_t)k * mx + a];  // Q[a][k]
      pxt[(size_t)a * mx + k] = q * swx[a];     // Px[k][a]
      bxt[(size_t)k * mx + a] = q / swx[a];     // Bx[a][k]
    }
  for (int j = 0; j < my; j++)
    for (int k = 0; k < mx; k++)
      inv_den[(size_t)j * mx + k] = 1.0 / (lx[k] + ly[j]);
  ST(green_checks_alloc(c, d));
  HIPCHK(c, hipMalloc((void **)&d.d_fd_mats, mats.size() * sizeof(double)));
  HIPCHK(c, hipMemcpy(d.d_fd_mats, mats.data(), mats.size() * sizeof(double),
                      hipMemcpyHostToDevice));
  HIPCHK(c, hipMalloc((void **)&d.d_fd_ix, (size_t)mx * sizeof(int)));
  HIPCHK(c, hipMalloc((void **)&d.d_fd_iy, (size_t)my * sizeof(int)));
  HIPCHK(c, hipMemcpy(d.d_fd_ix, ix.data(), (size_t)mx * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d.d_fd_iy, iy.data(), (size_t)my * sizeof(int), hipMemcpyHostToDevice));
  d.fd.mx = mx;
  d.fd.my = my;
  d.fd.ix = d.d_fd_ix;
  d.fd.iy = d.d_fd_iy;
  d.fd.py = d.d_fd_mats;
  d.fd.by = d.fd.py + nyy;
  d.fd.pxt = d.fd.by + nyy;
  d.fd.bxt = d.fd.pxt + nxx;
  d.fd.inv_den = d.fd.bxt + nxx;
  d.fd.t = d.d_fd_mats + (mats.size() - nyx);
  d.fd_state = 1;
  return EMCGPU_OK;
}

// corrections per step: the separable inverse is exact up to the hole term
// (one correction meets the stopping rule), the FP32 dense one takes two
inline int green_iters(const DeviceSim &d) { return d.fd_use ? 1 : kGreenIters; }

// grids whose contact nodes copy an inner node's field (the E-field kernel
// over all CUs forms it directly)
inline bool split_efield(const DeviceSim &d) { return d.ext[0] >= 3 && d.ext[1] >= 3; }

// arguments of the per-step Poisson / E-field launch (slot 1: the first
// step, which sets the contacts' applied voltages)
template <typename R>
DevFieldArgs<R> step_field_args(const emcgpu_ctx *c, const DeviceSim &d, bool reset_bc) {
  DevFieldArgs<R> f = field_args<R>(c, d);
  f.equilibrium = 0;
  f.reset_bc = reset_bc ? 1 : 0;
  f.do_poisson = 1;
  f.do_efield = 1;
  f.record = 1;
  f.nrpart = d.d_nrpart;
  f.interval = d.interval;
  // the E-field and force grid of a step over all CUs (k_dev_efield)
  // grids of at least 3 x 3 take the E-field over all CUs: from the
  // corrections' last check when the corrections run, else from its own
  // launch after the field kernel
  if (split_efield(d) && !(d.green_use && !reset_bc))
    f.do_efield = 0;
  if (split_efield(d) && d.green_use && !reset_bc)
    f.efield_if_solved = 1;
  if (d.green_use && !reset_bc) {
    f.green_iters = green_iters(d);
    f.green_flag = d.d_green_flag;
    f.green_acc = d.d_green_acc;
  }
  return f;
}

// the kernel sequence of one step (emcSimulation.hpp:164-193)
template <typename R>
emcgpu_status launch_step(emcgpu_ctx *c, DeviceSim &d, bool reset_bc) {
  const bool corr = d.green_use && !reset_bc;
  if (corr) {
    GreenArgs ga = green_args(d);
    if (split_efield(d)) {
      const DevFieldArgs<R> f = field_args<R>(c, d);
      ga.fuse_efield = 1;
      ga.iters = green_iters(d);
      ga.pm = f.pm;
      ga.acc_fp64 = sizeof(R) == sizeof(double);
      ga.vt = f.vt;
      ga.hxm = f.hxm;
      ga.hym = f.hym;
      ga.acc_scale = f.acc_scale;
      ga.bcx = f.bcx;
      ga.bcy = f.bcy;
      ga.ex = f.ex;
      ga.ey = f.ey;
      ga.force = f.acc;
    }
    HIPCHK(c, dev_launch_green(ga, green_iters(d), d.green_blocks, c->stream));
  }
  HIPCHK(c, dev_launch_field<R>(fargs_slot<R>(d, reset_bc ? 1 : 0), d.mg_lds, d.mg_global,
                                 c->stream));
  if (split_efield(d) && !corr)
    HIPCHK(c, dev_launch_efield<R>(fargs_slot<R>(d, reset_bc ? 1 : 0), (int)d.nn, c->stream));
  const DevContactArgs<R> ca = contact_args<R>(c, d);
  HIPCHK(c, dev_launch_step<R>(step_args<R>(c, d), ca, d.blocks, d.n_seg, d.step_lds, c->stream));
  if (ca.n_slots > 0)
    HIPCHK(c, dev_launch_contacts<R>(ca, d.max_inj, c->stream));
  HIPCHK(c, dev_launch_reduce<R>(ca, 1, c->stream));
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status compact(emcgpu_ctx *c, DeviceSim &d, long long n_total) {
  const int dst = 1 - d.cur;
  HIPCHK(c, dev_launch_compact<R>(parts<R>(d, d.cur), parts<R>(d, dst), d.d_blk, n_total,
                                  d.d_state, c->stream));
  d.cur = dst;
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status run_device_steps(emcgpu_ctx *c, DeviceSim &d, const emcgpu_device_step_args *s) {
  ST(need_model(c, s->dt));
  ST(need_regions(c, d));
  if (s->poisson_interval < 0)
    return fail(c, EMCGPU_ERR_ARG, "poisson_interval must be >= 0");
  d.interval = s->poisson_interval > 1 ? s->poisson_interval : 1;
  d.chan = s->channel_length > 0 ? 1 : 0;
  if (d.chan) {
    if (!(s->channel_x1 >= s->channel_x0))
      return fail(c, EMCGPU_ERR_ARG, "channel region: x1 must be >= x0");
    d.chan_x0 = s->channel_x0 * kNm;
    d.chan_x1 = s->channel_x1 * kNm;
    // getChannelDriftCurrent (emcBasicParticleHandler.hpp:224-238): q *
    // carriers per particle * sum v_x / length (electrons: charge -q is not
    // applied, as in the reference)
    d.chan_scale = phys::q / s->channel_length;
  } else {
    d.chan_x0 = d.chan_x1 = d.chan_scale = 0;
  }
  d.accumulate = s->accumulate ? 1 : 0;
  if (s->poisson_mode != EMCGPU_POISSON_AUTO && s->poisson_mode != EMCGPU_POISSON_MULTIGRID &&
      s->poisson_mode != EMCGPU_POISSON_DENSE)
    return fail(c, EMCGPU_ERR_ARG,
                "poisson_mode must be EMCGPU_POISSON_AUTO, _MULTIGRID or _DENSE");
  if (s->poisson_mode == EMCGPU_POISSON_AUTO)
    ST(fd_build(c, d));
  d.fd_use = (s->poisson_mode == EMCGPU_POISSON_AUTO && d.fd_state == 1) ? 1 : 0;
  if (s->poisson_mode == EMCGPU_POISSON_DENSE ||
      (s->poisson_mode == EMCGPU_POISSON_AUTO && !d.fd_use))
    ST(green_build(c, d));
  d.green_use = d.fd_use || (s->poisson_mode != EMCGPU_POISSON_MULTIGRID && d.green_state == 1);
  ST(ensure_hist(c, d, s->n_steps));
  const size_t nc = std::max(d.nContacts, 1);
  HIPCHK(c, dev_launch_zero_i(d.d_hexits, (long long)s->n_steps * nc, c->stream));
  HIPCHK(c, dev_launch_zero_i(d.d_hdels, (long long)s->n_steps * nc, c->stream));
  HIPCHK(c, dev_launch_zero_i(d.d_hinjs, (long long)s->n_steps * nc, c->stream));
  HIPCHK(c, dev_launch_mark(d.d_state, c->stream));
  ST(put_fargs<R>(c, d, 0, step_field_args<R>(c, d, false)));
  ST(put_fargs<R>(c, d, 1, step_field_args<R>(c, d, true)));
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  int done = 0;
  while (done < s->n_steps) {
    DevState st{};
    ST(read_state(c, d, st));
    if (st.overflow)
      return fail(c, EMCGPU_ERR_NOMEM, "device ensemble overflowed its capacity");
    const int k = std::min(kChunk, s->n_steps - done);
    const long long holes = st.n_total - st.n_alive;
    if (holes > (long long)(kHoleFrac * (double)st.n_alive) ||
        st.n_total + (long long)k * d.max_inj > d.capacity) {
      ST(compact<R>(c, d, st.n_total));
      ST(read_state(c, d, st));
      ST(ensure_capacity(c, d, st.n_total + (long long)kChunk * d.max_inj, st.n_total));
    }
    if (!d.bc_applied) {
      ST(launch_step<R>(c, d, true));
      d.bc_applied = true;
      done += 1;
      continue;
    }
    if (s->use_graph && k == kChunk) {
      if (!d.gexec || d.graph_cur != d.cur || d.graph_cap != d.capacity ||
          d.graph_chan != d.chan || d.graph_acc != d.accumulate ||
          d.graph_green != d.green_use + d.fd_use || d.graph_interval != d.interval ||
          d.graph_x0 != d.chan_x0 ||
          d.graph_x1 != d.chan_x1 || d.graph_scale != d.chan_scale) {
        if (d.gexec) {
          (void)hipGraphExecDestroy(d.gexec);
          d.gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        emcgpu_status cs = EMCGPU_OK;
        for (int q = 0; q < k && cs == EMCGPU_OK; q++)
          cs = launch_step<R>(c, d, false);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (cs)
          return cs;
        if (e != hipSuccess)
          return fail(c, EMCGPU_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&d.gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess)
          return fail(c, EMCGPU_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
        d.graph_cur = d.cur;
        d.graph_cap = d.capacity;
        d.graph_chan = d.chan;
        d.graph_acc = d.accumulate;
        d.graph_green = d.green_use + d.fd_use;
        d.graph_interval = d.interval;
        d.graph_x0 = d.chan_x0;
        d.graph_x1 = d.chan_x1;
        d.graph_scale = d.chan_scale;
      }
      HIPCHK(c, hipGraphLaunch(d.gexec, c->stream));
    } else {
      for (int q = 0; q < k; q++)
        ST(launch_step<R>(c, d, false));
    }
    done += k;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess)
    c->ms_total += ms;
  c->launches += s->n_steps;
  DevState st{};
  ST(read_state(c, d, st));
  if (st.overflow)
    return fail(c, EMCGPU_ERR_NOMEM, "device ensemble overflowed its capacity");
  const size_t ns = s->n_steps;
  auto get = [&](void *dst, const void *src, size_t bytes) -> emcgpu_status {
    if (dst)
      HIPCHK(c, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return EMCGPU_OK;
  };
  const size_t ncb = (size_t)d.nContacts;
  ST(get(s->exits, d.d_hexits, ns * ncb * sizeof(int)));
  ST(get(s->deleted, d.d_hdels, ns * ncb * sizeof(int)));
  ST(get(s->injected, d.d_hinjs, ns * ncb * sizeof(int)));
  ST(get(s->n_particles, d.d_hnpart, ns * sizeof(long long)));
  ST(get(s->poisson_iters, d.d_hiters, ns * sizeof(int)));
  ST(get(s->poisson_resid, d.d_hresid, ns * sizeof(double)));
  ST(get(s->checksum, d.d_hcheck, ns * sizeof(double)));
  if (d.chan)
    ST(get(s->channel_current, d.d_hchan, ns * sizeof(double)));
  return EMCGPU_OK;
}

// host Philox uniform of the device streams (the kernels' stream_uniform)
double host_uniform(uint32_t j, uint32_t step, uint32_t pid, uint32_t dom, uint64_t seed) {
  return stream_uniform<double>(j, step, pid, dom, (uint32_t)seed, (uint32_t)(seed >> 32));
}

template <typename R>
emcgpu_status device_init(emcgpu_ctx *c, DeviceSim &d, uint64_t seed, size_t *n_out) {
  ST(need_model(c, c->params_dt > 0 ? c->params_dt : 1e-15));
  ST(need_regions(c, d));
  d.seed = seed;
  // equilibrium potential (calcEquilibriumCharacteristics,
  // emcSimulation.hpp:139-146) from the charge-neutral guess asinh(C/2)
  // (initPotential, emcSimulationResults.hpp:208-213)
  std::vector<double> pot0(d.nn);
  for (int k = 0; k < d.nn; k++)
    pot0[k] = std::asinh(0.5 * d.dopn[k]);
  HIPCHK(c, hipMemcpy(d.d_pot, pot0.data(), sizeof(double) * d.nn, hipMemcpyHostToDevice));
  DevFieldArgs<R> f = field_args<R>(c, d);
  f.equilibrium = 1;
  f.reset_bc = 1;
  f.do_poisson = 1;
  f.do_efield = 1;
  f.max_iter = std::max(d.max_iter, 200);
  f.iters_out = d.d_iters;
  f.resid_out = d.d_resid;
  ST(run_field<R>(c, d, f));
  // particles per node (getInitialNrParticles, emcElectron.hpp:48-61, and the
  // remainder rule of generateInitialParticles,
  // emcAbstractParticleHandler.hpp:133-148): the whole carriers, then one
  // more with probability equal to the remainder (a Philox draw per node);
  // the density is the doping, or exp(equilibrium potential) * n_i
  std::vector<double> eq;
  if (d.init_pot) {
    eq.resize(d.nn);
    // the context stream is non-blocking: wait for the solve before copying
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(eq.data(), d.d_pot, sizeof(double) * d.nn, hipMemcpyDeviceToHost));
  }
  std::vector<long long> prefix(d.nn + 1, 0);
  for (int iy = 0; iy < d.ext[1]; iy++)
    for (int ix = 0; ix < d.ext[0]; ix++) {
      const int k = idx(d, ix, iy);
      double dens = d.init_pot ? std::exp(eq[k]) * d.ni : d.doping[k];
      if (ix == 0 || ix == d.ext[0] - 1)
        dens *= 0.5;
      if (iy == 0 || iy == d.ext[1] - 1)
        dens *= 0.5;
      double nr = dens * d.cellVolume;
      long long n = 0;
      while (nr >= 1) {
        n++;
        nr -= 1.0;
      }
      if (host_uniform(0, 0xFFFFFFFEu, (uint32_t)k, kDomCount, seed) < nr)
        n++;
      prefix[k + 1] = prefix[k] + n;
    }
  const long long n0 = prefix[d.nn];
  ST(ensure_capacity(c, d, n0 + (long long)kChunk * d.max_inj + 1, 0));
  d.cur = 0;
  long long *d_prefix = nullptr;
  HIPCHK(c, hipMalloc((void **)&d_prefix, sizeof(long long) * prefix.size()));
  HIPCHK(c, hipMemcpy(d_prefix, prefix.data(), sizeof(long long) * prefix.size(),
                      hipMemcpyHostToDevice));
  DevState s{};
  s.n_total = n0;
  s.n_alive = n0;
  s.next_pid = n0;
  s.capacity = d.capacity;
  ST(write_state(c, d, s));
  d.chan = 0;
  d.accumulate = 0;
  HIPCHK(c, hipMemset(d.d_avg_pot, 0, sizeof(double) * d.nn));
  HIPCHK(c, hipMemset(d.d_avg_conc, 0, sizeof(double) * d.nn));
  HIPCHK(c, hipMemset(d.d_chan_corr, 0, sizeof(double)));
  const DevContactArgs<R> ca = contact_args<R>(c, d);
  hipError_t e = dev_launch_init<R>(ca, d_prefix, n0, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_prefix);
  if (e != hipSuccess)
    return fail(c, EMCGPU_ERR_HIP, "initial particles: %s", hipGetErrorString(e));
  // first charge assignment (emcSimulation.hpp:150-160)
  HIPCHK(c, dev_launch_zero(d.d_corr, d.nn, c->stream));
  if (!d.lds_dep)
    HIPCHK(c, dev_launch_zero_r<R>((R *)d.d_dep, d.nn, c->stream));
  DevStepArgs<R> sa = step_args<R>(c, d);
  sa.move = 0;
  sa.hist = DevHist{};
  HIPCHK(c, dev_launch_step<R>(sa, ca, d.blocks, d.n_seg, d.step_lds, c->stream));
  HIPCHK(c, dev_launch_reduce<R>(ca, 0, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d.initialized = true;
  d.bc_applied = false;
  if (n_out)
    *n_out = (size_t)n0;
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status set_particles_t(emcgpu_ctx *c, DeviceSim &d, size_t n, const double *pos,
                              const double *k, const double *E, const double *tau,
                              const int32_t *valley, const int32_t *sub) {
  ST(ensure_capacity(c, d, (long long)n + (long long)kChunk * d.max_inj + 1, 0));
  d.cur = 0;
  std::vector<R> buf(n);
  auto put = [&](void *dst, auto fn) -> emcgpu_status {
    for (size_t i = 0; i < n; i++)
      buf[i] = (R)fn(i);
    if (n)
      HIPCHK(c, hipMemcpy(dst, buf.data(), n * sizeof(R), hipMemcpyHostToDevice));
    return EMCGPU_OK;
  };
  ST(put(d.px[0], [&](size_t i) { return pos[2 * i] * kNm; }));
  ST(put(d.py[0], [&](size_t i) { return pos[2 * i + 1] * kNm; }));
  ST(put(d.pkx[0], [&](size_t i) { return k[3 * i] * 1e-9; }));
  ST(put(d.pky[0], [&](size_t i) { return k[3 * i + 1] * 1e-9; }));
  ST(put(d.pkz[0], [&](size_t i) { return k[3 * i + 2] * 1e-9; }));
  ST(put(d.pE[0], [&](size_t i) { return E[i]; }));
  ST(put(d.ptau[0], [&](size_t i) { return tau[i] * 1e15; }));
  std::vector<uint32_t> vs(n), pid(n);
  for (size_t i = 0; i < n; i++) {
    vs[i] = ((uint32_t)valley[i] << 16) | (uint32_t)sub[i];
    pid[i] = (uint32_t)i;
  }
  if (n) {
    HIPCHK(c, hipMemcpy(d.pvs[0], vs.data(), n * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.ppid[0], pid.data(), n * 4, hipMemcpyHostToDevice));
  }
  DevState s{};
  ST(read_state(c, d, s));
  s.n_total = (long long)n;
  s.n_alive = (long long)n;
  s.next_pid = (long long)n;
  s.capacity = d.capacity;
  s.overflow = 0;
  ST(write_state(c, d, s));
  d.initialized = true;
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status get_particles_t(emcgpu_ctx *c, DeviceSim &d, size_t *n, double *pos, double *k,
                              double *E, double *tau, int32_t *valley, int32_t *sub) {
  DevState s{};
  ST(read_state(c, d, s));
  const long long nt = s.n_total;
  std::vector<uint32_t> vs(nt);
  if (nt)
    HIPCHK(c, hipMemcpy(vs.data(), d.pvs[d.cur], nt * 4, hipMemcpyDeviceToHost));
  std::vector<long long> live;
  for (long long i = 0; i < nt; i++)
    if (vs[i] != kHole)
      live.push_back(i);
  const size_t cap = *n;
  *n = live.size();
  if (live.size() > cap && (pos || k || E || tau || valley || sub))
    return fail(c, EMCGPU_ERR_ARG, "output holds %zu particles, the ensemble has %zu", cap,
                live.size());
  std::vector<R> b(nt);
  auto pull = [&](const void *src) -> emcgpu_status {
    if (nt)
      HIPCHK(c, hipMemcpy(b.data(), src, nt * sizeof(R), hipMemcpyDeviceToHost));
    return EMCGPU_OK;
  };
  if (pos) {
    ST(pull(d.px[d.cur]));
    for (size_t j = 0; j < live.size(); j++)
      pos[2 * j] = (double)b[live[j]] / kNm;
    ST(pull(d.py[d.cur]));
    for (size_t j = 0; j < live.size(); j++)
      pos[2 * j + 1] = (double)b[live[j]] / kNm;
  }
  if (k) {
    void *src[3] = {d.pkx[d.cur], d.pky[d.cur], d.pkz[d.cur]};
    for (int q = 0; q < 3; q++) {
      ST(pull(src[q]));
      for (size_t j = 0; j < live.size(); j++)
        k[3 * j + q] = (double)b[live[j]] * 1e9;
    }
  }
  if (E) {
    ST(pull(d.pE[d.cur]));
    for (size_t j = 0; j < live.size(); j++)
      E[j] = (double)b[live[j]];
  }
  if (tau) {
    ST(pull(d.ptau[d.cur]));
    for (size_t j = 0; j < live.size(); j++)
      tau[j] = (double)b[live[j]] * 1e-15;
  }
  for (size_t j = 0; j < live.size(); j++) {
    if (valley)
      valley[j] = (int32_t)(vs[live[j]] >> 16);
    if (sub)
      sub[j] = (int32_t)(vs[live[j]] & 0xffffu);
  }
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status set_efield_t(emcgpu_ctx *c, DeviceSim &d, const double *ex, const double *ey) {
  HIPCHK(c, hipMemcpy(d.d_ex, ex, sizeof(double) * d.nn, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d.d_ey, ey, sizeof(double) * d.nn, hipMemcpyHostToDevice));
  std::vector<R> acc(2 * (size_t)d.nn);
  for (int k = 0; k < d.nn; k++) {
    acc[2 * k] = (R)(ex[k] * kAccScale);
    acc[2 * k + 1] = (R)(ey[k] * kAccScale);
  }
  HIPCHK(c, hipMemcpy(d.d_acc, acc.data(), sizeof(R) * acc.size(), hipMemcpyHostToDevice));
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status drift_t(emcgpu_ctx *c, DeviceSim &d, const emcgpu_device_drift_args *s) {
  ST(need_model(c, s->dt));
  DevState st{};
  ST(read_state(c, d, st));
  const long long n = st.n_total;
  if (s->mode == EMCGPU_MODE_REPLAY) {
    if (c->precision != EMCGPU_FP64)
      return fail(c, EMCGPU_ERR_ARG, "replay mode needs an FP64 context");
    if (!s->draws || !s->cursor || s->draw_stride <= 0)
      return fail(c, EMCGPU_ERR_ARG, "replay needs draws, draw_stride and cursor");
  }
  std::vector<void *> tmp;
  auto dmal = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess)
      return nullptr;
    tmp.push_back(p);
    return p;
  };
  auto cleanup = [&]() {
    for (void *p : tmp)
      (void)hipFree(p);
  };
  DevStepArgs<R> a = step_args<R>(c, d);
  a.n = n;
  a.step = s->step0;
  a.n_steps = s->n_steps;
  a.key0 = (uint32_t)s->seed;
  a.key1 = (uint32_t)(s->seed >> 32);
  a.dt_s = s->dt;
  a.exit_step = (int *)dmal(n * sizeof(int));
  a.exit_contact = (int *)dmal(n * sizeof(int));
  a.draw_overflow = (int *)dmal(sizeof(int));
  if (!a.exit_step || !a.exit_contact || !a.draw_overflow) {
    cleanup();
    return fail(c, EMCGPU_ERR_NOMEM, "drift buffers");
  }
  std::vector<int> m1(n, -1);
  hipError_t e = hipMemcpy(a.exit_step, m1.data(), n * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(a.exit_contact, m1.data(), n * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemset(a.draw_overflow, 0, sizeof(int));
  const long long ns = s->n_steps;
  if (s->tr_pos) {
    a.tr_pos = (double *)dmal(ns * n * 2 * sizeof(double));
    a.tr_k = (double *)dmal(ns * n * 3 * sizeof(double));
    a.tr_E = (double *)dmal(ns * n * sizeof(double));
    a.tr_events = (int *)dmal(ns * n * sizeof(int));
  }
  if (s->mode == EMCGPU_MODE_REPLAY) {
    a.draws = (const uint64_t *)dmal(n * s->draw_stride * sizeof(uint64_t));
    a.cursor = (uint64_t *)dmal(n * sizeof(uint64_t));
    a.draw_stride = s->draw_stride;
    if (e == hipSuccess)
      e = hipMemcpy((void *)a.draws, s->draws, n * s->draw_stride * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(a.cursor, s->cursor, n * 8, hipMemcpyHostToDevice);
    if (s->grain_tau) {
      a.grain_tau = (double *)dmal(n * sizeof(double));
      if (e == hipSuccess)
        e = hipMemcpy(a.grain_tau, s->grain_tau, n * 8, hipMemcpyHostToDevice);
    }
  }
  if (e == hipSuccess && n > 0)
    e = dev_launch_drift<R>(a, s->mode, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  int over = 0;
  if (e == hipSuccess)
    e = hipMemcpy(&over, a.draw_overflow, sizeof(int), hipMemcpyDeviceToHost);
  auto back = [&](void *dst, const void *src, size_t bytes) {
    if (e == hipSuccess && dst && src)
      e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  };
  back(s->exit_step, a.exit_step, n * sizeof(int));
  back(s->exit_contact, a.exit_contact, n * sizeof(int));
  back(s->tr_pos, a.tr_pos, ns * n * 2 * sizeof(double));
  back(s->tr_k, a.tr_k, ns * n * 3 * sizeof(double));
  back(s->tr_E, a.tr_E, ns * n * sizeof(double));
  back(s->tr_events, a.tr_events, ns * n * sizeof(int));
  if (s->mode == EMCGPU_MODE_REPLAY) {
    back(s->cursor, a.cursor, n * 8);
    back(s->grain_tau, a.grain_tau, n * 8);
  }
  cleanup();
  if (e != hipSuccess)
    return fail(c, EMCGPU_ERR_HIP, "device drift: %s", hipGetErrorString(e));
  if (over)
    return fail(c, EMCGPU_ERR_DRAWS, "%d particles ran out of recorded draws", over);
  return EMCGPU_OK;
}

// the contact rule of handleOhmicContacts on caller positions: the step
// kernel's classification (no motion), the contact scan and the deletions
template <typename R>
emcgpu_status contacts_t(emcgpu_ctx *c, DeviceSim &d, const double *pos, size_t n, int *keep,
                         int *inject) {
  ST(need_model(c, c->params_dt > 0 ? c->params_dt : 1e-15));
  // the caller's positions become the ensemble (no k-space state needed)
  std::vector<double> k(3 * n, 0.0), E(n, 0.0), tau(n, 0.0);
  std::vector<int32_t> z(n, 0);
  ST(set_particles_t<R>(c, d, n, pos, k.data(), E.data(), tau.data(), z.data(), z.data()));
  ST(ensure_hist(c, d, 1));
  const size_t nc = std::max(d.nContacts, 1);
  HIPCHK(c, dev_launch_zero_i(d.d_hdels, nc, c->stream));
  HIPCHK(c, dev_launch_zero_i(d.d_hinjs, nc, c->stream));
  HIPCHK(c, dev_launch_zero_i(d.d_hexits, nc, c->stream));
  HIPCHK(c, dev_launch_mark(d.d_state, c->stream));
  DevStepArgs<R> sa = step_args<R>(c, d);
  sa.move = 0;
  const DevContactArgs<R> ca = contact_args<R>(c, d);
  HIPCHK(c, dev_launch_step<R>(sa, ca, d.blocks, d.n_seg, d.step_lds, c->stream));
  if (ca.n_slots > 0)
    HIPCHK(c, dev_launch_contacts<R>(ca, 0, c->stream));
  HIPCHK(c, dev_launch_zero(d.d_corr, d.nn, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<uint32_t> vs(n);
  if (n)
    HIPCHK(c, hipMemcpy(vs.data(), d.pvs[d.cur], n * 4, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++)
    keep[i] = vs[i] != kHole ? 1 : 0;
  std::vector<int> inj(std::max<size_t>(d.slot_node.size(), 1), 0);
  if (!d.slot_node.empty())
    HIPCHK(c, hipMemcpy(inj.data(), d.d_slot_inj, d.slot_node.size() * sizeof(int),
                        hipMemcpyDeviceToHost));
  for (int q = 0; q < d.nn; q++)
    inject[q] = 0;
  for (size_t s = 0; s < d.slot_node.size(); s++)
    inject[d.slot_node[s]] = inj[s];
  // the ensemble now holds the caller's positions only
  d.initialized = false;
  return EMCGPU_OK;
}

}  // namespace
}  // namespace host
}  // namespace emcgpu

using namespace emcgpu;
using namespace emcgpu::host;

#define DISPATCH(c, fn, ...) \
  ((c)->precision == EMCGPU_FP32 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

extern "C" {

emcgpu_status emcgpu_device_setup(emcgpu_ctx *c, const emcgpu_device_desc *desc) {
  if (!c || !desc)
    return c ? fail(c, EMCGPU_ERR_ARG, "desc must be non-NULL") : EMCGPU_ERR_ARG;
  if (c->track_pos)
    return fail(c, EMCGPU_ERR_STATE, "a bulk context with a periodic box cannot hold a device");
  DeviceSim *d = new DeviceSim();
  emcgpu_status st = build_geometry(c, *d, *desc);
  if (st == EMCGPU_OK) {
    derive_nodes(*d);
    if ((int)d->slot_node.size() > kMaxSlots)
      st = fail(c, EMCGPU_ERR_ARG, "%zu ohmic nodes exceed %d", d->slot_node.size(), kMaxSlots);
  }
  if (st == EMCGPU_OK)
    st = build_multigrid(c, *d);
  if (st == EMCGPU_OK && c->device >= 0) {
    if (d->slot_node.empty())  // the contact and injection kernels need ohmic nodes
      st = fail(c, EMCGPU_ERR_ARG, "the device needs at least one ohmic contact");
    else
      st = upload_device(c, *d);
  }
  if (st != EMCGPU_OK) {
    device_sim_destroy(d);
    return st;
  }
  device_sim_destroy(c->dsim);
  c->dsim = d;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_device_info(const emcgpu_ctx *c, int *ext, double *scal) {
  if (!c || !c->dsim)
    return EMCGPU_ERR_STATE;
  const DeviceSim &d = *c->dsim;
  if (ext) {
    ext[0] = d.ext[0];
    ext[1] = d.ext[1];
  }
  if (scal) {
    scal[0] = d.Vt;
    scal[1] = d.debye;
    scal[2] = d.cellVolume;
    scal[3] = d.normCellVolume;
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_device_grids(const emcgpu_ctx *c, double *doping, int *region,
                                  int *contact) {
  if (!c || !c->dsim)
    return EMCGPU_ERR_STATE;
  const DeviceSim &d = *c->dsim;
  for (int k = 0; k < d.nn; k++) {
    if (doping)
      doping[k] = d.doping[k];
    if (region)
      region[k] = d.region[k];
    if (contact)
      contact[k] = d.cnode[k];
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_poisson(emcgpu_ctx *c, double *pot, const double *conc, int equilibrium,
                             int reset_bc, int *iters, double *resid) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  if (!pot || (!equilibrium && !conc))
    return fail(c, EMCGPU_ERR_ARG, "pot (and conc out of equilibrium) must be non-NULL");
  DeviceSim &d = *c->dsim;
  HIPCHK(c, hipMemcpy(d.d_pot_t, pot, sizeof(double) * d.nn, hipMemcpyHostToDevice));
  if (conc)
    HIPCHK(c, hipMemcpy(d.d_conc_t, conc, sizeof(double) * d.nn, hipMemcpyHostToDevice));
  auto go = [&](auto tag) -> emcgpu_status {
    using R = decltype(tag);
    DevFieldArgs<R> f = field_args<R>(c, d);
    f.pot = d.d_pot_t;
    f.conc = d.d_conc_t;
    f.equilibrium = equilibrium ? 1 : 0;
    f.reset_bc = reset_bc ? 1 : 0;
    f.do_poisson = 1;
    f.iters_out = d.d_iters;
    f.resid_out = d.d_resid;
    if (equilibrium)
      f.max_iter = std::max(d.max_iter, 200);
    return run_field<R>(c, d, f);
  };
  st = c->precision == EMCGPU_FP32 ? go(0.0f) : go(0.0);
  if (st)
    return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(pot, d.d_pot_t, sizeof(double) * d.nn, hipMemcpyDeviceToHost));
  if (iters)
    HIPCHK(c, hipMemcpy(iters, d.d_iters, sizeof(int), hipMemcpyDeviceToHost));
  if (resid)
    HIPCHK(c, hipMemcpy(resid, d.d_resid, sizeof(double), hipMemcpyDeviceToHost));
  return EMCGPU_OK;
}

emcgpu_status emcgpu_efield(emcgpu_ctx *c, const double *pot, double *ex, double *ey) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  if (!pot || !ex || !ey)
    return fail(c, EMCGPU_ERR_ARG, "pot, ex and ey must be non-NULL");
  DeviceSim &d = *c->dsim;
  HIPCHK(c, hipMemcpy(d.d_pot_t, pot, sizeof(double) * d.nn, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(d.d_ex_t, 0, sizeof(double) * d.nn));
  HIPCHK(c, hipMemset(d.d_ey_t, 0, sizeof(double) * d.nn));
  auto go = [&](auto tag) -> emcgpu_status {
    using R = decltype(tag);
    DevFieldArgs<R> f = field_args<R>(c, d);
    f.pot = d.d_pot_t;
    f.ex = d.d_ex_t;
    f.ey = d.d_ey_t;
    f.acc = nullptr;
    f.do_efield = 1;
    return run_field<R>(c, d, f);
  };
  st = c->precision == EMCGPU_FP32 ? go(0.0f) : go(0.0);
  if (st)
    return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(ex, d.d_ex_t, sizeof(double) * d.nn, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ey, d.d_ey_t, sizeof(double) * d.nn, hipMemcpyDeviceToHost));
  return EMCGPU_OK;
}

emcgpu_status emcgpu_deposit(emcgpu_ctx *c, const double *pos, size_t n, double *grid) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  if ((n && !pos) || !grid)
    return fail(c, EMCGPU_ERR_ARG, "pos and grid must be non-NULL");
  DeviceSim &d = *c->dsim;
  double *dpos = nullptr;
  HIPCHK(c, hipMalloc((void **)&dpos, std::max<size_t>(n, 1) * 2 * sizeof(double)));
  hipError_t e = hipMemcpy(dpos, pos, n * 2 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d.d_conc_t, grid, sizeof(double) * d.nn, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = dev_launch_assign_pos(geom_m(d, nullptr), dpos, (long long)n, d.d_conc_t, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess)
    e = hipMemcpy(grid, d.d_conc_t, sizeof(double) * d.nn, hipMemcpyDeviceToHost);
  (void)hipFree(dpos);
  if (e != hipSuccess)
    return fail(c, EMCGPU_ERR_HIP, "deposit: %s", hipGetErrorString(e));
  return EMCGPU_OK;
}

emcgpu_status emcgpu_interpolate(emcgpu_ctx *c, const double *ex, const double *ey,
                                 const double *pos, size_t n, double charge, double *force) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  if (!ex || !ey || (n && (!pos || !force)))
    return fail(c, EMCGPU_ERR_ARG, "ex, ey, pos and force must be non-NULL");
  DeviceSim &d = *c->dsim;
  std::vector<double> e2(2 * (size_t)d.nn);
  for (int k = 0; k < d.nn; k++) {
    e2[2 * k] = ex[k];
    e2[2 * k + 1] = ey[k];
  }
  double *buf = nullptr;
  HIPCHK(c, hipMalloc((void **)&buf, std::max<size_t>(n, 1) * 4 * sizeof(double)));
  double *dpos = buf, *dforce = buf + 2 * std::max<size_t>(n, 1);
  hipError_t e = hipMemcpy(d.d_ex_t, e2.data(), e2.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dpos, pos, n * 2 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = dev_launch_interp_pos(geom_m(d, d.d_ex_t), dpos, (long long)n, charge, dforce, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess)
    e = hipMemcpy(force, dforce, n * 2 * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess)
    return fail(c, EMCGPU_ERR_HIP, "interpolate: %s", hipGetErrorString(e));
  return EMCGPU_OK;
}

// ---- 3-D particle-mesh operators --------------------------------------
namespace {
emcgpu_status pm3_grid(emcgpu_ctx *c, const emcgpu_grid3 *g, int min_nodes, PMGrid3 &out,
                       size_t &nn) {
  if (!c)
    return EMCGPU_ERR_ARG;
  if (!g)
    return fail(c, EMCGPU_ERR_ARG, "grid must be non-NULL");
  nn = 1;
  for (int d = 0; d < 3; d++) {
    if (g->ext[d] < min_nodes || g->ext[d] > (1 << 20))
      return fail(c, EMCGPU_ERR_ARG, "ext[%d] = %d must be in [%d, 2^20]", d, g->ext[d],
                  min_nodes);
    if (!(g->spacing[d] > 0.0) || !std::isfinite(g->spacing[d]))
      return fail(c, EMCGPU_ERR_ARG, "spacing[%d] must be positive", d);
    nn *= (size_t)g->ext[d];
  }
  if (nn > ((size_t)1 << 31))
    return fail(c, EMCGPU_ERR_ARG, "a 3-D grid holds at most 2^31 nodes");
  if (g->pm_scheme != EMCGPU_PM_NGP && g->pm_scheme != EMCGPU_PM_CIC)
    return fail(c, EMCGPU_ERR_ARG,
                "3-D schemes are EMCGPU_PM_NGP and EMCGPU_PM_CIC (the reference's NEC is 2-D)");
  out = PMGrid3{g->ext[0], g->ext[1], g->ext[2], g->spacing[0], g->spacing[1], g->spacing[2],
                g->pm_scheme};
  return EMCGPU_OK;
}
// after the arguments are checked (so a host-only context reports them)
emcgpu_status pm3_gpu(emcgpu_ctx *c) {
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "the 3-D particle-mesh operators need a GPU context");
  return EMCGPU_OK;
}
}  // namespace

emcgpu_status emcgpu_deposit3(emcgpu_ctx *c, const emcgpu_grid3 *g3, const double *pos, size_t n,
                              double *grid) {
  PMGrid3 g;
  size_t nn = 0;
  emcgpu_status st = pm3_grid(c, g3, 1, g, nn);
  if (st)
    return st;
  if ((n && !pos) || !grid)
    return fail(c, EMCGPU_ERR_ARG, "pos and grid must be non-NULL");
  if ((st = pm3_gpu(c)))
    return st;
  double *buf = nullptr;
  HIPCHK(c, hipMalloc((void **)&buf, (3 * n + nn) * sizeof(double)));
  double *dpos = buf, *dgrid = buf + 3 * n;
  hipError_t e = hipMemcpy(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dgrid, grid, nn * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = dev_launch_pm3_assign(g, dpos, (long long)n, dgrid, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess)
    e = hipMemcpy(grid, dgrid, nn * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess)
    return fail(c, EMCGPU_ERR_HIP, "deposit3: %s", hipGetErrorString(e));
  return EMCGPU_OK;
}

emcgpu_status emcgpu_interpolate3(emcgpu_ctx *c, const emcgpu_grid3 *g3, const double *ex,
                                  const double *ey, const double *ez, const double *pos, size_t n,
                                  double charge, double *force) {
  PMGrid3 g;
  size_t nn = 0;
  emcgpu_status st = pm3_grid(c, g3, 1, g, nn);
  if (st)
    return st;
  if (!ex || !ey || !ez || (n && (!pos || !force)))
    return fail(c, EMCGPU_ERR_ARG, "ex, ey, ez, pos and force must be non-NULL");
  if ((st = pm3_gpu(c)))
    return st;
  double *buf = nullptr;
  HIPCHK(c, hipMalloc((void **)&buf, (3 * nn + 6 * n) * sizeof(double)));
  double *de = buf, *dpos = buf + 3 * nn, *dforce = dpos + 3 * n;
  hipError_t e = hipMemcpy(de, ex, nn * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(de + nn, ey, nn * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(de + 2 * nn, ez, nn * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = dev_launch_pm3_interp(g, de, de + nn, de + 2 * nn, dpos, (long long)n, charge, dforce,
                              c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess)
    e = hipMemcpy(force, dforce, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess)
    return fail(c, EMCGPU_ERR_HIP, "interpolate3: %s", hipGetErrorString(e));
  return EMCGPU_OK;
}

emcgpu_status emcgpu_efield3(emcgpu_ctx *c, const emcgpu_grid3 *g3, double vt,
                             const signed char *bc, const double *pot, double *ex, double *ey,
                             double *ez) {
  PMGrid3 g;
  size_t nn = 0;
  emcgpu_status st = pm3_grid(c, g3, 3, g, nn);
  if (st)
    return st;
  if (!bc || !pot || !ex || !ey || !ez)
    return fail(c, EMCGPU_ERR_ARG, "bc, pot, ex, ey and ez must be non-NULL");
  if (!(vt > 0.0))
    return fail(c, EMCGPU_ERR_ARG, "vt must be positive");
  // every node on a face normal to d needs a status (-1 or a contact),
  // every other node -2
  for (int d = 0; d < 3; d++)
    for (size_t m = 0; m < nn; m++) {
      const size_t nxy = (size_t)g3->ext[0] * g3->ext[1];
      const size_t cd = d == 0 ? m % g3->ext[0] : d == 1 ? (m % nxy) / g3->ext[0] : m / nxy;
      const bool face = cd == 0 || cd == (size_t)g3->ext[d] - 1;
      const int b = bc[d * nn + m];
      if (face ? b < -1 : b != -2)
        return fail(c, EMCGPU_ERR_ARG,
                    "bc[%d][%zu] = %d: a node on a face normal to dimension %d needs -1 or a "
                    "contact index, any other node -2",
                    d, m, b, d);
    }
  if ((st = pm3_gpu(c)))
    return st;
  double *buf = nullptr;
  HIPCHK(c, hipMalloc((void **)&buf, 4 * nn * sizeof(double) + 3 * nn));
  double *dpot = buf, *de = buf + nn;
  signed char *dbc = (signed char *)(buf + 4 * nn);
  hipError_t e = hipMemcpy(dpot, pot, nn * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dbc, bc, 3 * nn, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemset(de, 0, 3 * nn * sizeof(double));
  if (e == hipSuccess)
    e = dev_launch_pm3_efield(g, vt, dbc, dpot, de, de + nn, de + 2 * nn, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  double *outs[3] = {ex, ey, ez};
  for (int d = 0; d < 3 && e == hipSuccess; d++)
    e = hipMemcpy(outs[d], de + d * nn, nn * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess)
    return fail(c, EMCGPU_ERR_HIP, "efield3: %s", hipGetErrorString(e));
  return EMCGPU_OK;
}

emcgpu_status emcgpu_contacts(emcgpu_ctx *c, const double *pos, size_t n, int *keep,
                              int *inject) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  if ((n && (!pos || !keep)) || !inject)
    return fail(c, EMCGPU_ERR_ARG, "pos, keep and inject must be non-NULL");
  return DISPATCH(c, contacts_t, c, *c->dsim, pos, n, keep, inject);
}

emcgpu_status emcgpu_device_set_particles(emcgpu_ctx *c, size_t n, const double *pos,
                                          const double *k, const double *energy,
                                          const double *tau, const int32_t *valley,
                                          const int32_t *sub) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  if (n && (!pos || !k || !energy || !tau || !valley || !sub))
    return fail(c, EMCGPU_ERR_ARG, "particle arrays must be non-NULL");
  const Instance &in = c->inst[0];
  for (size_t i = 0; i < n; i++) {
    if (valley[i] < 0 || valley[i] >= (int)in.valleys.size())
      return fail(c, EMCGPU_ERR_ARG, "particle %zu: valley %d out of range", i, valley[i]);
    if (sub[i] < 0 || sub[i] >= in.valleys[valley[i]].degeneracy)
      return fail(c, EMCGPU_ERR_ARG, "particle %zu: subvalley %d out of range", i, sub[i]);
  }
  return DISPATCH(c, set_particles_t, c, *c->dsim, n, pos, k, energy, tau, valley, sub);
}

emcgpu_status emcgpu_device_get_particles(const emcgpu_ctx *c0, size_t *n, double *pos,
                                          double *k, double *energy, double *tau,
                                          int32_t *valley, int32_t *sub) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  if (!n)
    return fail(c, EMCGPU_ERR_ARG, "n must be non-NULL");
  return DISPATCH(c, get_particles_t, c, *c->dsim, n, pos, k, energy, tau, valley, sub);
}

emcgpu_status emcgpu_device_set_efield(emcgpu_ctx *c, const double *ex, const double *ey) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  if (!ex || !ey)
    return fail(c, EMCGPU_ERR_ARG, "ex and ey must be non-NULL");
  return DISPATCH(c, set_efield_t, c, *c->dsim, ex, ey);
}

emcgpu_status emcgpu_device_drift(emcgpu_ctx *c, const emcgpu_device_drift_args *args) {
  emcgpu_status st = need_device(c, true);
  if (st)
    return st;
  if (!args || args->n_steps < 0 || !(args->dt > 0))
    return fail(c, EMCGPU_ERR_ARG, "invalid drift arguments");
  return DISPATCH(c, drift_t, c, *c->dsim, args);
}

emcgpu_status emcgpu_device_init(emcgpu_ctx *c, uint64_t seed, size_t *n_particles) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  return DISPATCH(c, device_init, c, *c->dsim, seed, n_particles);
}

emcgpu_status emcgpu_device_step(emcgpu_ctx *c, const emcgpu_device_step_args *args) {
  emcgpu_status st = need_device(c, true);
  if (st)
    return st;
  if (!args || args->n_steps < 0 || !(args->dt > 0))
    return fail(c, EMCGPU_ERR_ARG, "invalid step arguments");
  if (args->n_steps == 0)
    return EMCGPU_OK;
  return DISPATCH(c, run_device_steps, c, *c->dsim, args);
}

emcgpu_status emcgpu_device_fields(const emcgpu_ctx *c0, double *pot, double *ex, double *ey,
                                   double *conc, double *nr_part) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  DeviceSim &d = *c->dsim;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t b = sizeof(double) * d.nn;
  if (pot)
    HIPCHK(c, hipMemcpy(pot, d.d_pot, b, hipMemcpyDeviceToHost));
  if (ex)
    HIPCHK(c, hipMemcpy(ex, d.d_ex, b, hipMemcpyDeviceToHost));
  if (ey)
    HIPCHK(c, hipMemcpy(ey, d.d_ey, b, hipMemcpyDeviceToHost));
  if (conc)
    HIPCHK(c, hipMemcpy(conc, d.d_conc, b, hipMemcpyDeviceToHost));
  if (nr_part)
    HIPCHK(c, hipMemcpy(nr_part, d.d_nrpart, b, hipMemcpyDeviceToHost));
  return EMCGPU_OK;
}

emcgpu_status emcgpu_device_averages(emcgpu_ctx *c, double *pot, double *conc,
                                     long long *n_steps, int reset) {
  emcgpu_status st = need_device(c);
  if (st)
    return st;
  DeviceSim &d = *c->dsim;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  DevState s{};
  ST(read_state(c, d, s));
  if (n_steps)
    *n_steps = s.avg_n;
  std::vector<double> buf(d.nn);
  for (int q = 0; q < 2; q++) {
    double *out = q == 0 ? pot : conc;
    if (!out)
      continue;
    HIPCHK(c, hipMemcpy(buf.data(), q == 0 ? d.d_avg_pot : d.d_avg_conc, sizeof(double) * d.nn,
                        hipMemcpyDeviceToHost));
    // writeFinalResults (emcSimulationResults.hpp:181-183): sum / steps
    for (int k = 0; k < d.nn; k++)
      out[k] = s.avg_n > 0 ? buf[k] / (double)s.avg_n : 0.0;
  }
  if (reset) {
    HIPCHK(c, hipMemset(d.d_avg_pot, 0, sizeof(double) * d.nn));
    HIPCHK(c, hipMemset(d.d_avg_conc, 0, sizeof(double) * d.nn));
    s.avg_n = 0;
    ST(write_state(c, d, s));
  }
  return EMCGPU_OK;
}

}  // extern "C"
