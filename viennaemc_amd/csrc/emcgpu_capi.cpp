// emcgpu_capi.cpp -- C ABI of the MI355X EMC particle step (host side).
//
// Owns the device memory of a context (SoA particle arena, per-instance
// parameter blocks, scatter tables, observable buffers), builds scatter
// tables in FP64 on the host exactly as emcScatterHandler does
// (include/emcScatterHandler.hpp:220-288), lowers valleys and mechanisms to
// the kernels' parameter blocks in scaled units, and drives the step,
// reduction and initialisation kernels on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/emcgpu.h"
#include "../include/emcgpu/emcgpuPhysics.hpp"
#include "emcgpu_device.h"

namespace emcgpu {
template <typename R>
hipError_t launch_steps(const StepArgs<R> &a, int mode, int ppt, hipStream_t st);
template <typename R>
hipError_t launch_init(const StepArgs<R> &a, double init_mean, double init_sigma, hipStream_t st);
template <typename R>
hipError_t launch_observe(const StepArgs<R> &a, hipStream_t st);
template <typename R>
hipError_t launch_bath_step(const BathStepArgs<R> &a, int n_groups, hipStream_t st);
template <typename R>
hipError_t launch_rebuild(const BathStepArgs<R> &a, int n_jobs, hipStream_t st);
hipError_t launch_rates_desc(const RateDescJob *jobs, int n_jobs, int n_levels, double dE,
                             double *out, hipStream_t st);
int reduce_chunks(long long max_count, int tile_size);
hipError_t launch_reduce(const double *partial, int n_steps, int n_tiles, int n_obs, int tile_size,
                         const long long *inst_start, const long long *inst_count, int n_inst,
                         int n_chunks, double *chunk_buf, double *out, int out_step0,
                         int out_total, hipStream_t st);
}  // namespace emcgpu

using namespace emcgpu;

#include "emcgpu_host.h"

using namespace emcgpu::host;

namespace {

// ---- static scratch of emcgpu_step_noalloc ---------------------------------
// The per-step partials and the reduction pairs of the allocation-free step
// live in this file-scope array: each GPU that loads the library holds its
// own copy (64 MiB) until the process ends.  In host code the name is only a
// stand-in; the device address of each GPU's copy is looked up once per
// device (scratch_setup, from emcgpu_create with that device current) and
// only that address is handed to kernels.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __attribute__((used)) __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

struct ScratchTable {
  std::mutex mu;
  std::vector<double *> ptr;       // per device ordinal
  std::vector<std::string> error;  // per device: why the lookup failed
};
ScratchTable &scratch_table() {
  static ScratchTable t;
  return t;
}

// the calling thread's current device must be `device`
void scratch_setup(int device) {
  ScratchTable &t = scratch_table();
  std::lock_guard<std::mutex> lock(t.mu);
  if ((int)t.ptr.size() <= device) {
    t.ptr.resize(device + 1, nullptr);
    t.error.resize(device + 1);
  }
  if (t.ptr[device])
    return;
  double *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    t.error[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  t.ptr[device] = p;
}

// the device address of `device`'s scratch, or null with the reason in *why
double *scratch_ptr(int device, std::string *why) {
  ScratchTable &t = scratch_table();
  std::lock_guard<std::mutex> lock(t.mu);
  if (device < 0 || device >= (int)t.ptr.size()) {
    *why = "no scratch lookup for this device";
    return nullptr;
  }
  if (!t.ptr[device])
    *why = t.error[device];
  return t.ptr[device];
}

// ---- scaled-unit lowering of an instance -------------------------------
// k [1/nm] = k [1/m] * 1e-9, t [fs] = t [s] * 1e15
constexpr double kToNm = 1e-9;
constexpr double kToFs = 1e15;

void normalize3(double d[3]) {
  double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (n == 0.0)
    return;
  for (int i = 0; i < 3; i++)
    d[i] /= n;
}

// ---- mechanism kinds ---------------------------------------------------
inline bool is_hot_kind(int k) {
  return k == EMCGPU_MECH_HOT_FROEHLICH_AB || k == EMCGPU_MECH_HOT_FROEHLICH_EM;
}
inline bool is_scr_kind(int k) {
  return k == EMCGPU_MECH_SCREENED_FROEHLICH_AB || k == EMCGPU_MECH_SCREENED_FROEHLICH_EM;
}
inline bool is_scr_hot_kind(int k) {
  return k == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB || k == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_EM;
}
inline bool reads_bath(int k) { return is_hot_kind(k) || is_scr_hot_kind(k); }
// a mechanism whose rate the GPU rebuilds during a run
inline bool dyn_mech(const Instance &in, const emcgpu_mech_desc &m) {
  return reads_bath(m.kind) || (in.plasma() && is_scr_kind(m.kind));
}
bool dyn_valley(const Instance &in, size_t v) {
  for (int j : in.table_mech[v])
    if (dyn_mech(in, in.mechs[j]))
      return true;
  return false;
}

// rebuildWindowSums + getNqInWindow (emcPhononBath.hpp:109-124, 378-394) on
// the host copy of a bath
double window_nq_host(const BathHost &b, double qs2, double mean, double qmin, double qmax) {
  const int nb = b.d.n_bins;
  if (nb < 2 || qmax <= qmin)
    return mean;
  long lo = (long)std::floor(qmin / b.d.dq), hi = (long)std::ceil(qmax / b.d.dq);
  if (lo < 0)
    lo = 0;
  if (hi > (long)nb)
    hi = (long)nb;
  if (hi - lo < 1)
    return mean;
  double cw = 0, cwn = 0, wlo = 0, wnlo = 0;
  for (long i = 0; i < hi; i++) {
    if (i == lo) {
      wlo = cw;
      wnlo = cwn;
    }
    const double q = ((double)(std::size_t)i + 0.5) * b.d.dq;
    const double denom = q * q + qs2;
    const double w = (denom > 0.0) ? q / denom : 0.0;
    cw = cw + w;
    cwn = cwn + w * b.nq[i];
  }
  const double wsum = cw - wlo;
  if (wsum <= 0.0)
    return mean;
  return (cwn - wnlo) / wsum;
}

template <typename R>
void lower_instance(const emcgpu_ctx *c, const Instance &in, double dt_s, InstDev<R> &P,
                    const std::vector<long long> &table_off) {
  std::memset(&P, 0, sizeof(P));
  P.n_valleys = (int)in.valleys.size();
  P.n_levels = in.n_levels;
  P.dE = (R)in.dE;
  P.inv_dE = (R)(1.0 / in.dE);
  P.dt = (R)(dt_s * kToFs);
  P.vt = (R)(phys::kB / phys::q * in.temperature);
  for (int d = 0; d < 3; d++)
    P.box[d] = (R)(in.box[d] * 1e9);
  for (int v = 0; v < P.n_valleys; v++) {
    const emcgpu_valley_desc &vd = in.valleys[v];
    const phys::ValleyMass &m = in.vm[v];
    P.tau0[v] = (R)(in.tau0[v] * kToFs);
    P.n_tables[v] = (int)in.table_mech[v].size();
    P.table_off[v] = table_off[v];
    ValleyDev<R> &V = P.valley[v];
    V.c2 = (R)(phys::hbar * phys::hbar * 1e18 / (2.0 * m.effMassCond * phys::q));
    V.alpha = (R)m.alpha;
    V.cv = (R)(phys::hbar * 1e9 / m.effMassCond);
    V.deg = vd.degeneracy;
    for (int s = 0; s < vd.degeneracy; s++) {
      // M = R^T diag(vogt) R with R rows = normalised subvalley axes
      double M[3][3];
      if (phys::isAniso(vd.kind)) {
        double rows[3][3];
        for (int r = 0; r < 3; r++) {
          for (int j = 0; j < 3; j++)
            rows[r][j] = vd.axes[s][3 * r + j];
          normalize3(rows[r]);
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int r = 0; r < 3; r++)
              acc += rows[r][i] * m.vogt[r] * rows[r][j];
            M[i][j] = acc;
          }
      } else {
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            M[i][j] = (i == j) ? 1.0 : 0.0;
      }
      for (int i = 0; i < 3; i++) {
        double a = 0, w = 0;
        for (int j = 0; j < 3; j++) {
          a += M[i][j] * in.force[j];
          w += M[i][j] * in.dir[j];
        }
        // dk/dt = M F / hbar  ->  [1/nm per fs]
        V.accel[s][i] = (R)(a / phys::hbar * kToNm / kToFs);
        V.wdir[s][i] = (R)w;
      }
      const double hv[6] = {M[0][0], M[0][1], M[0][2], M[1][1], M[1][2], M[2][2]};
      for (int i = 0; i < 6; i++)
        V.hv[s][i] = (R)hv[i];
    }
    for (int j = 0; j < P.n_tables[v]; j++) {
      const emcgpu_mech_desc &md = in.mechs[in.table_mech[v][j]];
      MechDev<R> &mm = P.mech[v][j];
      mm.final_valley = v;
      mm.z = 1;
      switch (md.kind) {
      case EMCGPU_MECH_ACOUSTIC:
        mm.final_kind = kFinalIsotropicElastic;
        break;
      case EMCGPU_MECH_ZERO_IV_AB:
      case EMCGPU_MECH_ZERO_IV_EM:
      case EMCGPU_MECH_FIRST_IV_AB:
      case EMCGPU_MECH_FIRST_IV_EM: {
        mm.final_kind = kFinalIsotropicInelastic;
        mm.final_valley = md.final_valley;
        mm.z = md.z;
        const double delta = in.vm[md.final_valley].bottom - in.vm[md.valley].bottom;
        const bool ab = md.kind == EMCGPU_MECH_ZERO_IV_AB || md.kind == EMCGPU_MECH_FIRST_IV_AB;
        mm.shift = ab ? (R)(md.phonon_energy - delta) : (R)(-(md.phonon_energy + delta));
        for (int s = 0; s < kMaxDeg; s++)
          for (int z = 0; z < kMaxZ; z++)
            mm.final_sub[s][z] = (uint8_t)md.final_sub[s][z];
        break;
      }
      case EMCGPU_MECH_SCREENED_FROEHLICH_AB:
      case EMCGPU_MECH_SCREENED_FROEHLICH_EM:
      case EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB:
      case EMCGPU_MECH_SCREENED_HOT_FROEHLICH_EM: {
        const bool ab = md.kind == EMCGPU_MECH_SCREENED_FROEHLICH_AB ||
                        md.kind == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB;
        mm.final_kind = kFinalScreenedFroehlich;
        mm.shift = ab ? (R)md.phonon_energy : (R)(-md.phonon_energy);
        if (is_scr_hot_kind(md.kind)) {
          mm.bath1 = md.bath + 1;
          mm.qres = (md.q_resolved && md.q_resolved_angle) ? 1 : 0;
        }
        break;
      }
      default: {
        const bool ab = md.kind == EMCGPU_MECH_FROEHLICH_AB || md.kind == EMCGPU_MECH_HOT_FROEHLICH_AB;
        mm.final_kind = kFinalFroehlich;
        mm.shift = ab ? (R)md.phonon_energy : (R)(-md.phonon_energy);
        if (md.kind == EMCGPU_MECH_HOT_FROEHLICH_AB || md.kind == EMCGPU_MECH_HOT_FROEHLICH_EM)
          mm.bath1 = md.bath + 1;
        break;
      }
      }
    }
  }
  P.qs2 = (R)(in.qs2 * kToNm * kToNm);
  for (const auto &m : in.mechs)
    P.win_lds |= (is_scr_hot_kind(m.kind) && m.q_resolved && m.q_resolved_angle) ? 1 : 0;
  if (!c->d_win)
    P.win_lds = 0;
  for (size_t b = 0; b < in.baths.size() && c->d_win; b++)
    P.bath_win[b] = (const R *)c->d_win + 2 * (in.bath_off[b] + in.bath_index0 + (long long)b);
  P.n_baths = (int)in.baths.size();
  P.total_bins = in.total_bins;
  P.gbin_off = in.gbin_off;
  int loff = 0;
  for (int b = 0; b < P.n_baths; b++) {
    P.bath_nb[b] = in.baths[b].d.n_bins;
    P.bath_loff[b] = loff;
    P.bath_dq[b] = (R)(in.baths[b].d.dq * kToNm);
    loff += in.baths[b].d.n_bins;
  }
}

// ---- phonon baths -----------------------------------------------------
// planck / planckTemp / getMeanNq (emcPhononBath.hpp:127-137, 461-470)
double planck(double energy_ev, double temp_k) {
  const double x = phys::q * energy_ev / (phys::kB * temp_k);
  return 1.0 / (std::exp(x) - 1.0);
}
double planck_temp(double energy_ev, double n) {
  if (n <= 0.0)
    return 0.0;
  return phys::q * energy_ev / (phys::kB * std::log(1.0 + 1.0 / n));
}
double bath_mean(const BathHost &b) {
  double sum_w = 0.0, sum_wn = 0.0;
  for (int i = 0; i < b.d.n_bins; i++) {
    const double q = ((double)(std::size_t)i + 0.5) * b.d.dq;
    const double w = q * q;
    sum_w += w;
    sum_wn += w * b.nq[i];
  }
  return (sum_w > 0.0) ? sum_wn / sum_w : b.n0;
}

// copies the evolving device state of an instance's baths to the host copy
emcgpu_status pull_baths(emcgpu_ctx *c, Instance &in) {
  if (!in.dev_bath || c->device < 0)
    return EMCGPU_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t b = 0; b < in.baths.size(); b++) {
    BathHost &B = in.baths[b];
    HIPCHK(c, hipMemcpy(B.nq.data(), c->d_nq + in.bath_off[b], sizeof(double) * B.d.n_bins,
                        hipMemcpyDeviceToHost));
    BathVar v;
    HIPCHK(c, hipMemcpy(&v, c->d_bath_var + in.bath_index0 + (long long)b, sizeof(BathVar),
                        hipMemcpyDeviceToHost));
    B.n_ac = v.n_ac;
    B.n_to = v.n_to;
  }
  return EMCGPU_OK;
}

// copies the q_s^2 that plasma screening evolves on the GPU to the host
emcgpu_status pull_screening(emcgpu_ctx *c) {
  bool any = false;
  for (const auto &in : c->inst)
    any |= in.dev_scr;
  if (!any || c->device < 0 || !c->d_scr)
    return EMCGPU_OK;
  std::vector<double> v(c->n_inst);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(v.data(), c->d_scr, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < c->n_inst; i++)
    if (c->inst[i].dev_scr)
      c->inst[i].qs2 = v[i];
  return EMCGPU_OK;
}

// (re)creates the device bath arrays of all instances from their host state
emcgpu_status upload_bath_state(emcgpu_ctx *c) {
  for (auto &in : c->inst) {
    emcgpu_status st = pull_baths(c, in);
    if (st)
      return st;
  }
  std::vector<BathDev> bd;
  std::vector<BathVar> bv;
  std::vector<double> nq, prof;
  int lds = 0;
  for (auto &in : c->inst) {
    in.bath_index0 = (long long)bd.size();
    in.gbin_off = (long long)nq.size();
    in.bath_off.clear();
    in.total_bins = 0;
    for (const BathHost &B : in.baths) {
      BathDev d{};
      d.n_bins = B.d.n_bins;
      d.has_ac = B.has_ac;
      d.has_ridley = B.has_ridley;
      d.has_profile = !B.profile.empty();
      d.off = (long long)nq.size();
      d.dq = B.d.dq;
      d.tau_lo = B.d.tau_lo;
      d.n0 = B.n0;
      d.vsim = B.d.volume;
      d.lo_e = B.d.phonon_energy;
      d.lat_t = B.d.lattice_temp;
      d.ac_e = B.d.ac_phonon_energy;
      d.tau_ac = B.d.tau_acoustic;
      d.n_ac_eq = B.n_ac_eq;
      d.w_r = B.w_r;
      d.to_e = B.d.to_phonon_energy;
      d.tau_to = B.d.tau_to;
      d.n_to_eq = B.n_to_eq;
      d.woff = 2 * (d.off + (long long)bd.size());
      d.sum_w = 0.0;
      for (int i = 0; i < B.d.n_bins; i++) {
        const double q = ((double)(std::size_t)i + 0.5) * B.d.dq;
        d.sum_w += q * q;
      }
      in.bath_off.push_back(d.off);
      bd.push_back(d);
      bv.push_back(BathVar{B.n_ac, B.n_to, bath_mean(B), 0.0});
      nq.insert(nq.end(), B.nq.begin(), B.nq.end());
      if (B.profile.empty())
        prof.insert(prof.end(), B.d.n_bins, 0.0);
      else
        prof.insert(prof.end(), B.profile.begin(), B.profile.end());
      in.total_bins += B.d.n_bins;
    }
    lds = std::max(lds, in.total_bins);
  }
  dfreet(c->d_baths);
  dfreet(c->d_nq);
  dfreet(c->d_tau_prof);
  dfreet(c->d_tally);
  dfreet(c->d_bath_var);
  dfreet(c->d_win64);
  dfree(c->d_win);
  c->win_elems = 2 * ((long long)nq.size() + (long long)bd.size());
  c->n_baths = (int)bd.size();
  c->lds_bins = lds;
  c->tally_bins = (long long)nq.size();
  if (c->n_baths) {
    HIPCHK(c, hipMalloc((void **)&c->d_baths, sizeof(BathDev) * bd.size()));
    HIPCHK(c, hipMalloc((void **)&c->d_nq, sizeof(double) * nq.size()));
    HIPCHK(c, hipMalloc((void **)&c->d_tau_prof, sizeof(double) * prof.size()));
    HIPCHK(c, hipMalloc((void **)&c->d_tally,
                        sizeof(unsigned long long) * 2 * kTallyCopies * nq.size()));
    HIPCHK(c, hipMalloc((void **)&c->d_bath_var, sizeof(BathVar) * bv.size()));
    HIPCHK(c, hipMalloc((void **)&c->d_win64, sizeof(double) * c->win_elems));
    HIPCHK(c, hipMalloc(&c->d_win, (c->precision == EMCGPU_FP32 ? sizeof(float) : sizeof(double)) *
                                       c->win_elems));
    HIPCHK(c, hipMemcpy(c->d_baths, bd.data(), sizeof(BathDev) * bd.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_nq, nq.data(), sizeof(double) * nq.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_tau_prof, prof.data(), sizeof(double) * prof.size(),
                        hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->d_tally, 0, sizeof(unsigned long long) * 2 * kTallyCopies * nq.size()));
    HIPCHK(c, hipMemcpy(c->d_bath_var, bv.data(), sizeof(BathVar) * bv.size(),
                        hipMemcpyHostToDevice));
  }
  for (auto &in : c->inst)
    in.dev_bath = !in.baths.empty();
  c->bath_layout_dirty = false;
  return EMCGPU_OK;
}

template <typename R>
BathStepArgs<R> bath_args(const emcgpu_ctx *c, double dt, double *hist, int hist_steps,
                          int hist_step, int rebuild) {
  BathStepArgs<R> a;
  std::memset(&a, 0, sizeof(a));
  a.groups = c->d_groups;
  a.baths = c->d_baths;
  a.nq = c->d_nq;
  a.tau_prof = c->d_tau_prof;
  a.tally = c->d_tally;
  a.tally_bins = c->tally_bins;
  a.var = c->d_bath_var;
  a.dt = dt;
  a.hist = hist;
  a.hist_steps = hist_steps;
  a.hist_step = hist_step;
  a.rebuild = rebuild;
  a.jobs = c->d_jobs;
  a.arena = c->d_rate_arena;
  a.cum = (R *)c->d_cum;
  a.cum64 = c->d_cum64;
  a.inst = (InstDev<R> *)c->d_params;
  a.tau0_64 = c->d_tau0_64;
  a.scr = c->d_scr;
  a.win64 = c->d_win64;
  a.win = (R *)c->d_win;
  a.n_obs = c->n_obs;
  a.n_jobs = c->n_jobs;
  a.rb_max = c->d_rb_max;
  return a;
}

// rebuild jobs, their FP64 rate data and the FP64 table copies, then one
// rebuild so the device tables follow the current bath state
template <typename R>
emcgpu_status upload_rebuild(emcgpu_ctx *c, size_t cum_elems) {
  std::vector<RebuildJob> jobs;
  std::vector<BathInstDev> groups;
  std::vector<double> arena;
  std::vector<double> scr(c->n_inst, 0.0);
  for (int i = 0; i < c->n_inst; i++) {
    const Instance &in = c->inst[i];
    scr[i] = in.qs2;
    if (in.baths.empty() && !in.plasma())
      continue;
    BathInstDev g{};
    g.bath0 = (int)in.bath_index0;
    g.n_baths = (int)in.baths.size();
    g.job0 = (int)jobs.size();
    g.inst = i;
    g.plasma = in.plasma() ? 1 : 0;
    g.n_valleys = (int)in.valleys.size();
    g.scr_eps = in.scr.eps_static;
    g.scr_vol = in.scr.volume;
    for (const auto &m : in.mechs)
      g.need_win |= (is_scr_hot_kind(m.kind) && m.q_resolved) ? 1 : 0;
    for (size_t v = 0; in.has_tables && v < in.valleys.size(); v++) {
      const auto &tm = in.table_mech[v];
      if (!dyn_valley(in, v))
        continue;
      RebuildJob J{};
      J.inst = i;
      J.every_step = in.plasma() ? 1 : 0;
      J.valley = (int)v;
      J.n_tables = (int)tm.size();
      J.n_levels = in.n_levels;
      J.cum_off = in.cum_off[v];
      for (int t = 0; t < J.n_tables; t++) {
        const emcgpu_mech_desc &md = in.mechs[tm[t]];
        J.rate_off[t] = (long long)arena.size();
        arena.insert(arena.end(), in.rate_data[tm[t]].begin(), in.rate_data[tm[t]].end());
        if (is_hot_kind(md.kind)) {
          J.kind[t] = md.kind == EMCGPU_MECH_HOT_FROEHLICH_AB ? kRateHotAb : kRateHotEm;
          J.bath[t] = (int)(in.bath_index0 + md.bath);
          J.C[t] = in.hot_C[tm[t]];
        } else if (dyn_mech(in, md)) {
          // screened kinds: kI, kF per level; the occupation a constant, the
          // bath mean or the bath window mean
          const bool ab = md.kind == EMCGPU_MECH_SCREENED_FROEHLICH_AB ||
                          md.kind == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB;
          J.kind[t] = ab ? kRateScrAb : kRateScrEm;
          J.C[t] = in.hot_C[tm[t]];
          if (is_scr_hot_kind(md.kind)) {
            J.bath[t] = (int)(in.bath_index0 + md.bath);
            J.nsrc[t] = md.q_resolved ? kNWindow : kNMean;
          } else {
            J.nsrc[t] = kNConst;
            J.nconst[t] = in.mech_n0[tm[t]];
          }
        } else {
          J.kind[t] = kRateConst;
        }
      }
      jobs.push_back(J);
    }
    g.n_jobs = (int)jobs.size() - g.job0;
    groups.push_back(g);
  }
  dfreet(c->d_jobs);
  dfreet(c->d_groups);
  dfreet(c->d_rate_arena);
  dfreet(c->d_scr);
  dfreet(c->d_rb_max);
  c->n_jobs = (int)jobs.size();
  c->n_groups = (int)groups.size();
  HIPCHK(c, hipMalloc((void **)&c->d_scr, sizeof(double) * scr.size()));
  HIPCHK(c, hipMemcpy(c->d_scr, scr.data(), sizeof(double) * scr.size(), hipMemcpyHostToDevice));
  for (auto &in : c->inst)
    in.dev_scr = in.plasma();
  if (!groups.empty()) {
    HIPCHK(c, hipMalloc((void **)&c->d_groups, sizeof(BathInstDev) * groups.size()));
    HIPCHK(c, hipMemcpy(c->d_groups, groups.data(), sizeof(BathInstDev) * groups.size(),
                        hipMemcpyHostToDevice));
  }
  if (jobs.empty()) {
    // window sums of the baths still follow their state and q_s^2
    if (!groups.empty())
      HIPCHK(c, launch_rebuild<R>(bath_args<R>(c, 0.0, nullptr, 0, 0, 1), c->n_groups, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EMCGPU_OK;
  }
  HIPCHK(c, hipMalloc((void **)&c->d_jobs, sizeof(RebuildJob) * jobs.size()));
  HIPCHK(c, hipMalloc((void **)&c->d_rb_max, sizeof(double) * kRebuildChunks * jobs.size()));
  HIPCHK(c, hipMalloc((void **)&c->d_rate_arena, sizeof(double) * std::max<size_t>(arena.size(), 1)));
  HIPCHK(c, hipMemcpy(c->d_jobs, jobs.data(), sizeof(RebuildJob) * jobs.size(),
                      hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_rate_arena, arena.data(), sizeof(double) * arena.size(),
                      hipMemcpyHostToDevice));
  if (cum_elems * sizeof(double) > c->cum64_bytes) {
    dfreet(c->d_cum64);
    HIPCHK(c, hipMalloc((void **)&c->d_cum64, cum_elems * sizeof(double)));
    c->cum64_bytes = cum_elems * sizeof(double);
  }
  if (!c->d_tau0_64)
    HIPCHK(c, hipMalloc((void **)&c->d_tau0_64, sizeof(double) * kMaxV * c->n_inst));
  HIPCHK(c, launch_rebuild<R>(bath_args<R>(c, 0.0, nullptr, 0, 0, 1), c->n_groups, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EMCGPU_OK;
}

// builds the parameter blocks and the table arena of every instance
template <typename R>
emcgpu_status upload_params(emcgpu_ctx *c, double dt_s) {
  if (!c->params_dirty && dt_s == c->params_dt)
    return EMCGPU_OK;
  {
    emcgpu_status st = pull_screening(c);
    if (st)
      return st;
  }
  if (c->bath_layout_dirty) {
    emcgpu_status st = upload_bath_state(c);
    if (st)
      return st;
  }
  std::vector<InstDev<R>> P(c->n_inst);
  std::vector<R> cum;
  for (int i = 0; i < c->n_inst; i++) {
    const Instance &in = c->inst[i];
    std::vector<long long> off(kMaxV, 0);
    if (!in.has_tables)
      continue;
    for (size_t v = 0; v < in.valleys.size(); v++) {
      off[v] = (long long)cum.size();
      const int nt = (int)in.table_mech[v].size();
      // device layout [level][table]
      for (int l = 0; l < in.n_levels; l++)
        for (int j = 0; j < nt; j++)
          cum.push_back((R)in.cum[v][(size_t)j * in.n_levels + l]);
    }
    c->inst[i].cum_off = off;
    lower_instance<R>(c, in, dt_s, P[i], off);
  }
  if (cum.empty())
    cum.push_back(R(0));
  size_t bytes = cum.size() * sizeof(R);
  if (bytes > c->cum_bytes) {
    dfree(c->d_cum);
    HIPCHK(c, hipMalloc(&c->d_cum, bytes));
    c->cum_bytes = bytes;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_cum, cum.data(), bytes, hipMemcpyHostToDevice, c->stream));
  if (!c->d_params)
    HIPCHK(c, hipMalloc(&c->d_params, sizeof(InstDev<R>) * c->n_inst));
  HIPCHK(c, hipMemcpyAsync(c->d_params, P.data(), sizeof(InstDev<R>) * c->n_inst,
                           hipMemcpyHostToDevice, c->stream));
  // the host vectors die here: wait for the copies
  HIPCHK(c, hipStreamSynchronize(c->stream));
  emcgpu_status st = upload_rebuild<R>(c, cum.size());
  if (st)
    return st;
  c->params_dirty = false;
  c->params_dt = dt_s;
  return EMCGPU_OK;
}

// particles per lane of the production kernel when the caller leaves it to
// the library (measured best on MI355X for the Si bulk workload)
int default_ppt(int) { return 4; }

// doubles of the per-step partials of one step (one per wave and value)
size_t partial_per_step(const emcgpu_ctx *c, int n_tiles) {
  return (size_t)std::max(n_tiles, 1) * kSubs * (size_t)c->n_obs;
}
// doubles of the first reduction pass's (sum, compensation) pairs per step
size_t chunks_per_step(const emcgpu_ctx *c, int n_chunks) {
  return (size_t)c->n_inst * (size_t)c->n_obs * (size_t)n_chunks * 2;
}

emcgpu_status ensure_partial(emcgpu_ctx *c, int k_steps, int n_tiles) {
  size_t bytes = (size_t)k_steps * partial_per_step(c, n_tiles) * sizeof(double);
  if (bytes > c->partial_bytes) {
    dfreet(c->d_partial);
    HIPCHK(c, hipMalloc((void **)&c->d_partial, bytes));
    c->partial_bytes = bytes;
  }
  return EMCGPU_OK;
}

emcgpu_status ensure_obs(emcgpu_ctx *c, int n_steps) {
  size_t bytes = (size_t)c->n_inst * n_steps * c->n_obs * sizeof(double);
  if (bytes > c->obs_bytes) {
    dfreet(c->d_obs);
    HIPCHK(c, hipMalloc((void **)&c->d_obs, bytes));
    c->obs_bytes = bytes;
  }
  return EMCGPU_OK;
}

template <typename R>
void fill_common(const emcgpu_ctx *c, StepArgs<R> &a) {
  std::memset(&a, 0, sizeof(a));
  a.kx = (R *)c->kx;
  a.ky = (R *)c->ky;
  a.kz = (R *)c->kz;
  a.E = (R *)c->E;
  a.tau = (R *)c->tau;
  a.vs = c->vs;
  if (c->track_pos) {
    a.px = (R *)c->px;
    a.py = (R *)c->py;
    a.pz = (R *)c->pz;
  }
  a.inst = (const InstDev<R> *)c->d_params;
  a.cum = (const R *)c->d_cum;
  a.unit_inst = c->d_unit_inst;
  a.inst_start = c->d_inst_start;
  a.inst_count = c->d_inst_count;
  a.n_obs = c->n_obs;
  a.max_v = c->max_v;
  a.max_deg = c->max_deg;
  for (const auto &in : c->inst)
    for (const auto &m : in.mechs)
      a.ext_kinds |= (is_scr_kind(m.kind) || is_scr_hot_kind(m.kind)) ? 1 : 0;
}

// buffers of emcgpu_step_noalloc: the per-step partials and the reduction
// pairs in the library's static scratch, the observables in the caller's
// device buffer
struct NoAllocBuffers {
  double *scratch;     // device address of the static scratch array
  size_t scratch_len;  // its length in doubles
  double *obs;         // caller's device buffer [instance][step][value]
  size_t obs_len;      // its length in doubles
};

// reserve_only: size (and upload) everything a call with these arguments
// needs, then return without launching
template <typename R>
emcgpu_status run_steps(emcgpu_ctx *c, const emcgpu_step_args *s,
                        const NoAllocBuffers *na = nullptr, bool reserve_only = false) {
  if (c->capacity == 0)
    return fail(c, EMCGPU_ERR_STATE, "no particles (call emcgpu_resize first)");
  for (int i = 0; i < c->n_inst; i++) {
    if (c->inst[i].count && !c->inst[i].has_tables)
      return fail(c, EMCGPU_ERR_STATE, "instance %d has particles but no scatter tables", i);
  }
  const bool replay = s->mode == EMCGPU_MODE_REPLAY;
  if (replay && c->precision != EMCGPU_FP64)
    return fail(c, EMCGPU_ERR_ARG, "replay mode needs an FP64 context");
  if (replay && (!s->draws || !s->cursor || s->draw_stride == 0))
    return fail(c, EMCGPU_ERR_ARG, "replay mode needs draws, cursor and draw_stride");
  const int ppl = s->particles_per_lane;
  if (ppl != 0 && ppl != 1 && ppl != 2 && ppl != 4 && ppl != 8)
    return fail(c, EMCGPU_ERR_ARG, "particles_per_lane must be 0, 1, 2, 4 or 8 (got %d)", ppl);
  if (!(s->dt > 0) || s->n_steps < 0)
    return fail(c, EMCGPU_ERR_ARG, "dt must be > 0 and n_steps >= 0");
  if (s->n_steps == 0)
    return EMCGPU_OK;
  if (na) {
    // everything that would allocate is refused: the model must be on the
    // device for this dt already, no replay / trace / graph, no baths or
    // plasma screening (their per-step buffers)
    if (replay || s->trace_valley || s->trace_sub || s->trace_events || s->trace_k ||
        s->trace_energy || s->use_graph)
      return fail(c, EMCGPU_ERR_ARG,
                  "emcgpu_step_noalloc: production mode only (no replay, trace or graph)");
    if (c->n_baths > 0)
      return fail(c, EMCGPU_ERR_STATE, "emcgpu_step_noalloc: phonon baths are not supported");
    for (const auto &in : c->inst)
      if (in.plasma())
        return fail(c, EMCGPU_ERR_STATE, "emcgpu_step_noalloc: plasma screening is not supported");
    if (c->params_dirty || !c->d_params || s->dt != c->params_dt)
      return fail(c, EMCGPU_ERR_STATE,
                  "emcgpu_step_noalloc: the model is not on the device for dt = %g s (run "
                  "emcgpu_step once with this dt after the last model, table or field change)",
                  s->dt);
  }
  emcgpu_status st = upload_params<R>(c, s->dt);
  if (st)
    return st;
  if (s->steps_per_launch < 0)
    return fail(c, EMCGPU_ERR_ARG, "steps_per_launch must be >= 0");
  int K = s->steps_per_launch > 0 ? s->steps_per_launch
          : c->tuned_spl > 0      ? c->tuned_spl
                                  : std::min(s->n_steps, 64);
  K = std::min(K, s->n_steps);
  // phonon baths couple the particles between steps: one step per launch,
  // then the bath update and the table rebuild
  // plasma screening likewise: q_s^2 and the screened tables follow the
  // carriers after every step
  const bool baths = c->n_baths > 0;
  bool plasma = false;
  for (const auto &in : c->inst)
    plasma |= in.plasma();
  const bool dyn = baths || plasma;
  if (dyn)
    K = 1;
  const int every = s->rebuild_every > 1 ? s->rebuild_every : 1;
  // steps whose per-step partials are reduced by one pair of reduction
  // launches (a launch's steps, or up to 64 one-step launches with baths)
  int Kr = dyn ? std::min(s->n_steps, 64) : K;
  if (baths && !s->bath_record_only) {
    // per bath and step: the mean occupation, then the 3-value energy record
    const size_t hb = 4 * sizeof(double) * (size_t)c->n_baths * (size_t)s->n_steps;
    if (hb > c->bath_hist_bytes) {
      dfreet(c->d_bath_hist);
      HIPCHK(c, hipMalloc((void **)&c->d_bath_hist, hb));
      c->bath_hist_bytes = hb;
    }
    if (!reserve_only)
      c->bath_hist_steps = s->n_steps;
  } else if (!reserve_only) {
    c->bath_hist_steps = 0;
  }
  // particles per lane of the production kernel (replay and tracked
  // positions: one per lane)
  if (c->track_pos && ppl > 1)
    return fail(c, EMCGPU_ERR_ARG, "tracked positions run one particle per lane (ppl %d)", ppl);
  const int ppt = (replay || c->track_pos) ? 1
                  : ppl                    ? ppl
                  : c->tuned_ppl           ? c->tuned_ppl
                                           : default_ppt(c->precision);
  const int tile_size = kTile * ppt;
  const int n_tiles = (int)(c->capacity / tile_size);
  long long max_count = 0;
  for (const auto &in : c->inst)
    max_count = std::max<long long>(max_count, (long long)in.count);
  const int n_chunks = reduce_chunks(max_count, tile_size);
  double *d_partial = nullptr, *d_chunks = nullptr, *d_obs = nullptr;
  if (na) {
    // as many steps per launch as the scratch holds partials for (results
    // do not depend on the launch shape)
    const size_t per_step = partial_per_step(c, n_tiles) + chunks_per_step(c, n_chunks);
    const size_t kmax = na->scratch_len / per_step;
    if (kmax < 1)
      return fail(c, EMCGPU_ERR_ARG,
                  "emcgpu_step_noalloc: %zu bytes of scratch per step needed, the library's "
                  "scratch holds %zu bytes (fewer particles or instances per context)",
                  per_step * sizeof(double), na->scratch_len * sizeof(double));
    if ((size_t)K > kmax)
      K = (int)kmax;
    Kr = K;
    const size_t need = (size_t)c->n_inst * (size_t)s->n_steps * (size_t)c->n_obs;
    if (!na->obs || na->obs_len < need)
      return fail(c, EMCGPU_ERR_ARG,
                  "emcgpu_step_noalloc: the observable buffer needs %zu doubles "
                  "(instances x steps x values), got %zu", need, na->obs ? na->obs_len : 0);
    d_partial = na->scratch;
    d_chunks = na->scratch + (size_t)Kr * partial_per_step(c, n_tiles);
    d_obs = na->obs;
  } else {
    if ((st = ensure_partial(c, Kr, n_tiles)) || (st = ensure_obs(c, s->n_steps)))
      return st;
    const size_t bytes = (size_t)Kr * chunks_per_step(c, n_chunks) * sizeof(double);
    if (bytes > c->chunk_bytes) {
      dfreet(c->d_chunks);
      HIPCHK(c, hipMalloc((void **)&c->d_chunks, bytes));
      c->chunk_bytes = bytes;
    }
    d_partial = c->d_partial;
    d_chunks = c->d_chunks;
    d_obs = c->d_obs;
    if (reserve_only)
      return EMCGPU_OK;
    c->obs_steps = s->n_steps;
  }

  StepArgs<R> a;
  fill_common<R>(c, a);
  a.n_tiles = n_tiles;
  a.key0 = (uint32_t)s->seed;
  a.key1 = (uint32_t)(s->seed >> 32);
  a.partial = d_partial;
  a.dt_s = s->dt;
  if (baths) {
    a.tally = c->d_tally;
    a.tally_bins = c->tally_bins;
    a.lds_bins = c->lds_bins;
    for (const auto &in : c->inst)
      for (const auto &m : in.mechs)
        if (is_scr_hot_kind(m.kind) && m.q_resolved && m.q_resolved_angle)
          a.lds_win = std::max(a.lds_win, 2 * (in.total_bins + (int)in.baths.size()));
  }
  const size_t n0 = c->inst[0].count;

  // replay and trace buffers (instance 0)
  uint64_t *d_draws = nullptr, *d_cursor = nullptr;
  double *d_grain = nullptr;
  int *d_over = nullptr;
  int32_t *d_tv = nullptr, *d_ts = nullptr, *d_te = nullptr;
  double *d_tk = nullptr, *d_tE = nullptr, *d_tp = nullptr;
  bool capturing = false;
  auto cleanup = [&]() {
    if (capturing) {  // abandon a partial capture so the stream is usable
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(c->stream, &g);
      if (g)
        (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      capturing = false;
    }
    dfreet(d_draws);
    dfreet(d_cursor);
    dfreet(d_grain);
    dfreet(d_over);
    dfreet(d_tv);
    dfreet(d_ts);
    dfreet(d_te);
    dfreet(d_tk);
    dfreet(d_tE);
    dfreet(d_tp);
  };
#define HIPCHK_CLEAN(x)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      cleanup();                                                                         \
      return fail(c, EMCGPU_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));        \
    }                                                                                    \
  } while (0)
  if (replay) {
    if (c->n_inst != 1)
      return fail(c, EMCGPU_ERR_ARG, "replay mode supports one instance");
    HIPCHK_CLEAN(hipMalloc((void **)&d_draws, sizeof(uint64_t) * n0 * s->draw_stride + 8));
    HIPCHK_CLEAN(hipMalloc((void **)&d_cursor, sizeof(uint64_t) * n0 + 8));
    HIPCHK_CLEAN(hipMalloc((void **)&d_over, sizeof(int)));
    HIPCHK_CLEAN(hipMemcpyAsync(d_draws, s->draws, sizeof(uint64_t) * n0 * s->draw_stride,
                                hipMemcpyHostToDevice, c->stream));
    HIPCHK_CLEAN(hipMemcpyAsync(d_cursor, s->cursor, sizeof(uint64_t) * n0,
                                hipMemcpyHostToDevice, c->stream));
    HIPCHK_CLEAN(hipMemsetAsync(d_over, 0, sizeof(int), c->stream));
    if (s->grain_tau) {
      HIPCHK_CLEAN(hipMalloc((void **)&d_grain, sizeof(double) * n0 + 8));
      HIPCHK_CLEAN(hipMemcpyAsync(d_grain, s->grain_tau, sizeof(double) * n0,
                                  hipMemcpyHostToDevice, c->stream));
    }
    a.draws = d_draws;
    a.draw_stride = (long long)s->draw_stride;
    a.cursor = d_cursor;
    a.grain_tau = d_grain;
    a.overflow = d_over;
  }
  const bool trace = s->trace_valley && s->trace_sub && s->trace_events && s->trace_k &&
                     s->trace_energy;
  const bool trace_pos = trace && s->trace_pos && c->track_pos;
  if (trace) {
    const size_t m = n0 * (size_t)s->n_steps + 1;
    HIPCHK_CLEAN(hipMalloc((void **)&d_tv, sizeof(int32_t) * m));
    HIPCHK_CLEAN(hipMalloc((void **)&d_ts, sizeof(int32_t) * m));
    HIPCHK_CLEAN(hipMalloc((void **)&d_te, sizeof(int32_t) * m));
    HIPCHK_CLEAN(hipMalloc((void **)&d_tk, sizeof(double) * 3 * m));
    HIPCHK_CLEAN(hipMalloc((void **)&d_tE, sizeof(double) * m));
    if (trace_pos)
      HIPCHK_CLEAN(hipMalloc((void **)&d_tp, sizeof(double) * 3 * m));
    a.tr_n = (long long)n0;
  }

  const bool graph = s->use_graph && !replay && !trace && !s->timing;
  if (graph) {
    HIPCHK_CLEAN(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    capturing = true;
  }
  for (int done = 0; done < s->n_steps; done += K) {
    const int k = std::min(K, s->n_steps - done);
    a.n_steps = k;
    a.partial = d_partial + (size_t)(done % Kr) * partial_per_step(c, n_tiles);
    a.step0 = s->step0 + (uint32_t)done;
    if (trace) {
      const long long o = (long long)done * (long long)n0;
      a.tr_valley = d_tv + o;
      a.tr_sub = d_ts + o;
      a.tr_events = d_te + o;
      a.tr_k = d_tk + 3 * o;
      a.tr_E = d_tE + o;
      a.tr_pos = trace_pos ? d_tp + 3 * o : nullptr;
    }
    if (s->timing)
      HIPCHK_CLEAN(hipEventRecord(c->ev0, c->stream));
    HIPCHK_CLEAN(launch_steps<R>(a, replay ? 1 : 0, ppt, c->stream));
    if (s->timing) {
      HIPCHK_CLEAN(hipEventRecord(c->ev1, c->stream));
      HIPCHK_CLEAN(hipEventSynchronize(c->ev1));
      float ms = 0;
      HIPCHK_CLEAN(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      c->ms_total += ms;
      c->launches += 1;
    }
    if (dyn && !s->bath_record_only) {
      // end of the step of the hot-phonon drivers (hotPhononGa2O3.cpp:274-281;
      // with screening hotCarrierMHP.cpp:693-703)
      const long long gstep = (long long)s->step0 + done + 1;  // 1-based global step
      BathStepArgs<R> ba =
          bath_args<R>(c, s->dt, c->d_bath_hist, s->n_steps, done, gstep % every == 0);
      ba.energy = c->d_bath_hist + (size_t)c->n_baths * (size_t)s->n_steps;
      ba.partial = a.partial;
      ba.n_tiles = n_tiles;
      ba.tile_size = tile_size;
      ba.inst_start = c->d_inst_start;
      ba.inst_count = c->d_inst_count;
      HIPCHK_CLEAN(launch_bath_step<R>(ba, c->n_groups, c->stream));
    }
    // per-step partials of a batch of Kr steps are reduced together
    const int in_batch = (done % Kr) + k;
    if (in_batch == Kr || done + k == s->n_steps)
      HIPCHK_CLEAN(launch_reduce(d_partial, in_batch, n_tiles, c->n_obs, tile_size,
                                 c->d_inst_start, c->d_inst_count, c->n_inst, n_chunks,
                                 d_chunks, d_obs, done + k - in_batch, s->n_steps,
                                 c->stream));
  }
  if (graph) {
    hipGraph_t g = nullptr;
    capturing = false;
    HIPCHK_CLEAN(hipStreamEndCapture(c->stream, &g));
    bool updated = false;
    if (c->gexec) {
      hipGraphNode_t err_node = nullptr;
      hipGraphExecUpdateResult res;
      updated = hipGraphExecUpdate(c->gexec, g, &err_node, &res) == hipSuccess &&
                res == hipGraphExecUpdateSuccess;
      if (!updated) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
      }
    }
    hipError_t e = hipSuccess;
    if (!updated)
      e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK_CLEAN(e);
    HIPCHK_CLEAN(hipGraphLaunch(c->gexec, c->stream));
  }

  if (replay) {
    int over = 0;
    HIPCHK_CLEAN(hipMemcpyAsync(s->cursor, d_cursor, sizeof(uint64_t) * n0,
                                hipMemcpyDeviceToHost, c->stream));
    HIPCHK_CLEAN(hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (s->grain_tau)
      HIPCHK_CLEAN(hipMemcpyAsync(s->grain_tau, d_grain, sizeof(double) * n0,
                                  hipMemcpyDeviceToHost, c->stream));
    HIPCHK_CLEAN(hipStreamSynchronize(c->stream));
    if (over) {
      cleanup();
      return fail(c, EMCGPU_ERR_DRAWS, "%d particles ran out of recorded draws", over);
    }
  }
  if (trace) {
    const size_t m = n0 * (size_t)s->n_steps;
    HIPCHK_CLEAN(hipMemcpyAsync(s->trace_valley, d_tv, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK_CLEAN(hipMemcpyAsync(s->trace_sub, d_ts, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK_CLEAN(hipMemcpyAsync(s->trace_events, d_te, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK_CLEAN(hipMemcpyAsync(s->trace_k, d_tk, sizeof(double) * 3 * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK_CLEAN(hipMemcpyAsync(s->trace_energy, d_tE, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    if (trace_pos)
      HIPCHK_CLEAN(hipMemcpyAsync(s->trace_pos, d_tp, sizeof(double) * 3 * m,
                                  hipMemcpyDeviceToHost, c->stream));
  }
  if (replay || trace)
    HIPCHK_CLEAN(hipStreamSynchronize(c->stream));
  if (trace) {
    // the kernels trace k in 1/nm and positions in nm
    const size_t m = 3 * n0 * (size_t)s->n_steps;
    for (size_t i = 0; i < m; i++)
      s->trace_k[i] /= kToNm;
    if (trace_pos)
      for (size_t i = 0; i < m; i++)
        s->trace_pos[i] *= 1e-9;
  }
  cleanup();
#undef HIPCHK_CLEAN
  return EMCGPU_OK;
}

// FP64 host tables: fillScatterTables + renormalizeTables
emcgpu_status finish_tables(emcgpu_ctx *c, Instance &in, int n_levels, double max_energy,
                            const std::vector<std::vector<double>> &raw /*per mech*/) {
  const int nv = (int)in.valleys.size();
  in.n_levels = n_levels;
  in.max_energy = max_energy;
  in.dE = max_energy / (double)(std::size_t)n_levels;  // emcScatterHandler.hpp:73-75
  in.cum.assign(nv, {});
  in.table_mech.assign(nv, {});
  in.tau0.assign(nv, 2e-15);  // default of getTau (emcScatterHandler.hpp:79-84)
  for (int j = 0; j < (int)in.mechs.size(); j++)
    in.table_mech[in.mechs[j].valley].push_back(j);
  for (int v = 0; v < nv; v++) {
    const int nt = (int)in.table_mech[v].size();
    if (nt == 0)
      continue;
    std::vector<double> t((size_t)nt * n_levels);
    for (int j = 0; j < nt; j++)
      for (int l = 0; l < n_levels; l++)
        t[(size_t)j * n_levels + l] = raw[in.table_mech[v][j]][l];
    for (int j = 1; j < nt; j++)
      for (int l = 0; l < n_levels; l++)
        t[(size_t)j * n_levels + l] = t[(size_t)j * n_levels + l] + t[(size_t)(j - 1) * n_levels + l];
    double mx = t[(size_t)(nt - 1) * n_levels];
    for (int l = 1; l < n_levels; l++)
      if (mx < t[(size_t)(nt - 1) * n_levels + l])
        mx = t[(size_t)(nt - 1) * n_levels + l];
    if (!(mx > 0) || !std::isfinite(mx))
      return fail(c, EMCGPU_ERR_ARG, "valley %d: maximum total scatter rate is %g", v, mx);
    for (double &x : t)
      x /= mx;
    in.tau0[v] = 1. / mx;
    in.cum[v] = std::move(t);
  }
  in.has_tables = true;
  c->params_dirty = true;
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status init_thermal(emcgpu_ctx *c, int instance, uint64_t seed, double mean = 0.0,
                           double sigma = 0.0) {
  const Instance &in = c->inst[instance];
  if (!in.has_tables)
    return fail(c, EMCGPU_ERR_STATE, "instance %d: build tables before initialising", instance);
  emcgpu_status st = upload_params<R>(c, 1e-15);
  if (st)
    return st;
  StepArgs<R> a;
  fill_common<R>(c, a);
  a.key0 = (uint32_t)seed;
  a.key1 = (uint32_t)(seed >> 32);
  // one instance: its tiles only
  a.tile_base = (int)(in.start / kTile);
  a.n_tiles = (int)((in.count + kTile - 1) / kTile);
  HIPCHK(c, launch_init<R>(a, mean, sigma, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EMCGPU_OK;
}

}  // namespace

namespace {
template <typename R>
emcgpu_status set_particles_t(emcgpu_ctx *c, int instance, const double *k, const double *E,
                              const double *tau, const int32_t *valley, const int32_t *sub) {
  const Instance &in = c->inst[instance];
  const size_t n = in.count;
  if (n == 0)
    return EMCGPU_OK;
  std::vector<R> hx(n), hy(n), hz(n), he(n), ht(n);
  std::vector<uint32_t> hv(n);
  for (size_t i = 0; i < n; i++) {
    const int v = valley[i], s = sub[i];
    if (v < 0 || v >= (int)in.valleys.size() || s < 0 || s >= in.valleys[v].degeneracy)
      return fail(c, EMCGPU_ERR_ARG, "particle %zu: invalid valley %d / subvalley %d", i, v, s);
    hx[i] = (R)(k[3 * i] * kToNm);
    hy[i] = (R)(k[3 * i + 1] * kToNm);
    hz[i] = (R)(k[3 * i + 2] * kToNm);
    he[i] = (R)E[i];
    ht[i] = (R)(tau[i] * kToFs);
    hv[i] = ((uint32_t)v << 16) | (uint32_t)s;
  }
  const size_t off = (size_t)in.start;
  HIPCHK(c, hipMemcpyAsync((R *)c->kx + off, hx.data(), n * sizeof(R), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync((R *)c->ky + off, hy.data(), n * sizeof(R), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync((R *)c->kz + off, hz.data(), n * sizeof(R), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync((R *)c->E + off, he.data(), n * sizeof(R), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync((R *)c->tau + off, ht.data(), n * sizeof(R), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->vs + off, hv.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status get_particles_t(const emcgpu_ctx *c0, int instance, double *k, double *E,
                              double *tau, int32_t *valley, int32_t *sub) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  const Instance &in = c->inst[instance];
  const size_t n = in.count;
  if (n == 0)
    return EMCGPU_OK;
  std::vector<R> hx(n), hy(n), hz(n), he(n), ht(n);
  std::vector<uint32_t> hv(n);
  const size_t off = (size_t)in.start;
  HIPCHK(c, hipMemcpyAsync(hx.data(), (R *)c->kx + off, n * sizeof(R), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hy.data(), (R *)c->ky + off, n * sizeof(R), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hz.data(), (R *)c->kz + off, n * sizeof(R), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(he.data(), (R *)c->E + off, n * sizeof(R), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(ht.data(), (R *)c->tau + off, n * sizeof(R), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hv.data(), c->vs + off, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < n; i++) {
    if (k) {
      k[3 * i] = (double)hx[i] / kToNm;
      k[3 * i + 1] = (double)hy[i] / kToNm;
      k[3 * i + 2] = (double)hz[i] / kToNm;
    }
    if (E)
      E[i] = (double)he[i];
    if (tau)
      tau[i] = (double)ht[i] / kToFs;
    if (valley)
      valley[i] = (int32_t)(hv[i] >> 16);
    if (sub)
      sub[i] = (int32_t)(hv[i] & 0xffffu);
  }
  return EMCGPU_OK;
}
}  // namespace

namespace {
// the position arrays of the arena (zeroed); capacity may be 0
emcgpu_status alloc_positions(emcgpu_ctx *c) {
  dfree(c->px);
  dfree(c->py);
  dfree(c->pz);
  const size_t n = (size_t)std::max<long long>(c->capacity, 1), rs = real_size(c);
  HIPCHK(c, hipMalloc(&c->px, n * rs));
  HIPCHK(c, hipMalloc(&c->py, n * rs));
  HIPCHK(c, hipMalloc(&c->pz, n * rs));
  HIPCHK(c, hipMemsetAsync(c->px, 0, n * rs, c->stream));
  HIPCHK(c, hipMemsetAsync(c->py, 0, n * rs, c->stream));
  HIPCHK(c, hipMemsetAsync(c->pz, 0, n * rs, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status positions_t(emcgpu_ctx *c, int instance, double *pos, bool get) {
  const Instance &in = c->inst[instance];
  const size_t n = in.count, off = (size_t)in.start;
  if (n == 0)
    return EMCGPU_OK;
  std::vector<R> h[3];
  void *dev[3] = {c->px, c->py, c->pz};
  for (int d = 0; d < 3; d++) {
    h[d].resize(n);
    if (!get)
      for (size_t i = 0; i < n; i++)
        h[d][i] = (R)(pos[3 * i + d] * 1e9);
    if (get)
      HIPCHK(c, hipMemcpyAsync(h[d].data(), (R *)dev[d] + off, n * sizeof(R),
                               hipMemcpyDeviceToHost, c->stream));
    else
      HIPCHK(c, hipMemcpyAsync((R *)dev[d] + off, h[d].data(), n * sizeof(R),
                               hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (get)
    for (size_t i = 0; i < n; i++)
      for (int d = 0; d < 3; d++)
        pos[3 * i + d] = (double)h[d][i] * 1e-9;
  return EMCGPU_OK;
}

template <typename R>
emcgpu_status observe_t(emcgpu_ctx *c, double *out) {
  emcgpu_status st = upload_params<R>(c, c->params_dt > 0 ? c->params_dt : 1e-15);
  if (st)
    return st;
  const int n_tiles = (int)(c->capacity / kTile);
  if ((st = ensure_partial(c, 1, n_tiles)))
    return st;
  long long max_count = 0;
  for (const auto &in : c->inst)
    max_count = std::max<long long>(max_count, (long long)in.count);
  const int n_chunks = reduce_chunks(max_count, kTile);
  const size_t cb = chunks_per_step(c, n_chunks) * sizeof(double);
  if (cb > c->chunk_bytes) {
    dfreet(c->d_chunks);
    HIPCHK(c, hipMalloc((void **)&c->d_chunks, cb));
    c->chunk_bytes = cb;
  }
  if (!c->d_obs0)
    HIPCHK(c, hipMalloc((void **)&c->d_obs0, sizeof(double) * c->n_inst * (2 + kMaxV * (3 + kMaxDeg))));
  StepArgs<R> a;
  fill_common<R>(c, a);
  a.n_tiles = n_tiles;
  a.partial = c->d_partial;
  HIPCHK(c, launch_observe<R>(a, c->stream));
  HIPCHK(c, launch_reduce(c->d_partial, 1, n_tiles, c->n_obs, kTile, c->d_inst_start,
                          c->d_inst_count, c->n_inst, n_chunks, c->d_chunks, c->d_obs0, 0, 1,
                          c->stream));
  HIPCHK(c, hipMemcpyAsync(out, c->d_obs0, sizeof(double) * c->n_inst * c->n_obs,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EMCGPU_OK;
}
}  // namespace

// ===========================================================================
// exported C ABI
// ===========================================================================
extern "C" {

emcgpu_status emcgpu_create(const emcgpu_config *cfg, emcgpu_ctx **out) {
  if (!cfg || !out)
    return fail(nullptr, EMCGPU_ERR_ARG, "cfg and out must be non-NULL");
  *out = nullptr;
  if (cfg->n_instances < 1 || cfg->n_instances > 4096)
    return fail(nullptr, EMCGPU_ERR_ARG, "n_instances must be in [1, 4096]");
  if (cfg->precision != EMCGPU_FP64 && cfg->precision != EMCGPU_FP32)
    return fail(nullptr, EMCGPU_ERR_ARG, "unknown precision %d", cfg->precision);
  if (cfg->device == -1) {
    // host-only context: model, tables and validation, no device memory
    emcgpu_ctx *c = new emcgpu_ctx();
    c->device = -1;
    c->precision = cfg->precision;
    c->n_inst = cfg->n_instances;
    c->inst.resize(c->n_inst);
    *out = c;
    return EMCGPU_OK;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(nullptr, EMCGPU_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(nullptr, EMCGPU_ERR_ARG, "device %d out of range [0, %d)", cfg->device, ndev);
  if ((e = hipSetDevice(cfg->device)) != hipSuccess)
    return fail(nullptr, EMCGPU_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  // the allocation-free step's scratch address on this device (once per
  // device; a failure is reported by emcgpu_step_noalloc)
  scratch_setup(cfg->device);
  emcgpu_ctx *c = new emcgpu_ctx();
  c->device = cfg->device;
  c->precision = cfg->precision;
  c->n_inst = cfg->n_instances;
  c->inst.resize(c->n_inst);
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
    delete c;
    return fail(nullptr, EMCGPU_ERR_HIP, "stream/event creation: %s", hipGetErrorString(e));
  }
  *out = c;
  return EMCGPU_OK;
}

void emcgpu_destroy(emcgpu_ctx *c) {
  if (!c)
    return;
  if (c->device < 0) {
    device_sim_destroy(c->dsim);
    delete c;
    return;
  }
  (void)hipSetDevice(c->device);
  if (c->stream)
    (void)hipStreamSynchronize(c->stream);
  device_sim_destroy(c->dsim);
  c->dsim = nullptr;
  dfree(c->kx);
  dfree(c->ky);
  dfree(c->kz);
  dfree(c->E);
  dfree(c->tau);
  dfreet(c->vs);
  dfreet(c->d_unit_inst);
  dfreet(c->d_inst_start);
  dfreet(c->d_inst_count);
  dfree(c->d_params);
  dfree(c->d_cum);
  dfreet(c->d_partial);
  dfreet(c->d_chunks);
  dfreet(c->d_obs);
  dfreet(c->d_obs0);
  dfree(c->px);
  dfree(c->py);
  dfree(c->pz);
  dfreet(c->d_baths);
  dfreet(c->d_nq);
  dfreet(c->d_tau_prof);
  dfreet(c->d_tally);
  dfreet(c->d_bath_var);
  dfreet(c->d_bath_hist);
  dfreet(c->d_jobs);
  dfreet(c->d_rate_arena);
  dfreet(c->d_cum64);
  dfreet(c->d_tau0_64);
  dfreet(c->d_groups);
  dfreet(c->d_scr);
  dfreet(c->d_rb_max);
  dfreet(c->d_win64);
  dfree(c->d_win);
  if (c->gexec)
    (void)hipGraphExecDestroy(c->gexec);
  if (c->ev0)
    (void)hipEventDestroy(c->ev0);
  if (c->ev1)
    (void)hipEventDestroy(c->ev1);
  if (c->stream)
    (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *emcgpu_last_error(const emcgpu_ctx *c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

emcgpu_status emcgpu_set_model(emcgpu_ctx *c, int instance, const emcgpu_valley_desc *valleys,
                               int n_valleys, const emcgpu_mech_desc *mechs, int n_mechs) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!valleys || n_valleys < 1 || n_valleys > EMCGPU_MAX_VALLEYS)
    return fail(c, EMCGPU_ERR_ARG, "n_valleys must be in [1, %d]", EMCGPU_MAX_VALLEYS);
  if (n_mechs < 0 || n_mechs > EMCGPU_MAX_MECH || (n_mechs && !mechs))
    return fail(c, EMCGPU_ERR_ARG, "n_mechs must be in [0, %d]", EMCGPU_MAX_MECH);
  for (int v = 0; v < n_valleys; v++) {
    const emcgpu_valley_desc &d = valleys[v];
    if (d.kind < 0 || d.kind > 3)
      return fail(c, EMCGPU_ERR_ARG, "valley %d: unknown kind %d", v, d.kind);
    // emcAbstractValley::check (emcAbstractValley.hpp:79-90)
    if (d.degeneracy < 1 || d.degeneracy > EMCGPU_MAX_DEG)
      return fail(c, EMCGPU_ERR_ARG, "valley %d: degeneracy must be in [1, %d]", v, EMCGPU_MAX_DEG);
    if (d.alpha < 0)
      return fail(c, EMCGPU_ERR_ARG, "valley %d: NonParabolicity of a valley has to be positive", v);
    if (!(d.particle_mass > 0) || !(d.rel_mass[0] > 0) ||
        (phys::isAniso(d.kind) && (!(d.rel_mass[1] > 0) || !(d.rel_mass[2] > 0))))
      return fail(c, EMCGPU_ERR_ARG, "valley %d: masses must be positive", v);
    if (phys::isAniso(d.kind)) {
      for (int s = 0; s < d.degeneracy; s++) {
        const double *a = d.axes[s];
        auto dot = [&](int i, int j) {
          return a[3 * i] * a[3 * j] + a[3 * i + 1] * a[3 * j + 1] + a[3 * i + 2] * a[3 * j + 2];
        };
        // checkOrthogonalBasis (emcNonParabolicAnistropValley.hpp:208-220)
        if (dot(0, 1) != 0 || dot(0, 2) != 0 || dot(1, 2) != 0)
          return fail(c, EMCGPU_ERR_ARG,
                      "valley %d subvalley %d: the given coordinate system is not orthogonal", v, s);
        if (dot(0, 0) == 0 || dot(1, 1) == 0 || dot(2, 2) == 0)
          return fail(c, EMCGPU_ERR_ARG, "valley %d subvalley %d: zero axis", v, s);
      }
    }
  }
  for (int j = 0; j < n_mechs; j++) {
    const emcgpu_mech_desc &m = mechs[j];
    if (m.kind < 0 || m.kind > EMCGPU_MECH_SCREENED_HOT_FROEHLICH_EM)
      return fail(c, EMCGPU_ERR_ARG, "mechanism %d: unknown kind %d", j, m.kind);
    if (reads_bath(m.kind) && (m.bath < 0 || m.bath >= EMCGPU_MAX_BATHS))
      return fail(c, EMCGPU_ERR_ARG, "mechanism %d: bath index %d out of range", j, m.bath);
    if (m.valley < 0 || m.valley >= n_valleys)
      return fail(c, EMCGPU_ERR_ARG, "mechanism %d: Used index for Valley is invalid", j);
    if (m.kind >= EMCGPU_MECH_ZERO_IV_AB && m.kind <= EMCGPU_MECH_FIRST_IV_EM) {
      // emcZeroOrderInterValleyScatterMechanism::check (:132-165)
      if (m.final_valley < 0 || m.final_valley >= n_valleys)
        return fail(c, EMCGPU_ERR_ARG, "mechanism %d: idxFinalValley %d is not valid", j, m.final_valley);
      if (m.z < 1 || m.z > EMCGPU_MAX_Z)
        return fail(c, EMCGPU_ERR_ARG, "mechanism %d: z must be in [1, %d]", j, EMCGPU_MAX_Z);
      const int fdeg = valleys[m.final_valley].degeneracy;
      for (int s = 0; s < valleys[m.valley].degeneracy; s++)
        for (int z = 0; z < m.z; z++)
          if (m.final_sub[s][z] < 0 || m.final_sub[s][z] >= fdeg)
            return fail(c, EMCGPU_ERR_ARG,
                        "mechanism %d: Used idx %d for valley of degeneracy %d is not valid", j,
                        m.final_sub[s][z], fdeg);
    }
    if (m.kind != EMCGPU_MECH_ACOUSTIC && !(m.phonon_energy > 0))
      return fail(c, EMCGPU_ERR_ARG, "mechanism %d: phonon energy must be > 0", j);
  }
  Instance &in = c->inst[instance];
  in.valleys.assign(valleys, valleys + n_valleys);
  in.mechs.assign(mechs, mechs + n_mechs);
  in.vm.clear();
  for (const auto &d : in.valleys)
    in.vm.emplace_back(d.kind, d.rel_mass, d.particle_mass, d.alpha, d.bottom_energy);
  in.has_tables = false;
  in.has_hot = false;
  for (const auto &m : in.mechs)
    in.has_hot |= reads_bath(m.kind);
  // observable layout covers the largest model of the context
  int mv = 1, md = 1;
  for (const auto &i2 : c->inst) {
    mv = std::max(mv, (int)i2.valleys.size());
    for (const auto &d : i2.valleys)
      md = std::max(md, d.degeneracy);
  }
  c->max_v = mv;
  c->max_deg = md;
  c->n_obs = 2 + mv * (3 + md);
  c->params_dirty = true;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_set_rate_tables(emcgpu_ctx *c, int instance, int n_levels,
                                     double max_energy, const double *rates) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  Instance &in = c->inst[instance];
  if (in.valleys.empty())
    return fail(c, EMCGPU_ERR_STATE, "instance %d: set the model first", instance);
  if (n_levels < 1 || n_levels > EMCGPU_MAX_LEVELS || !(max_energy > 0) || !rates)
    return fail(c, EMCGPU_ERR_ARG, "n_levels in [1, %d], max_energy > 0 and rates required",
                EMCGPU_MAX_LEVELS);
  if (in.has_hot)
    return fail(c, EMCGPU_ERR_ARG,
                "hot-phonon rates follow their bath: build them with emcgpu_build_tables");
  if (in.plasma())
    for (const auto &m : in.mechs)
      if (is_scr_kind(m.kind))
        return fail(c, EMCGPU_ERR_ARG,
                    "plasma-screened rates follow the carriers: build them with emcgpu_build_tables");
  std::vector<std::vector<double>> raw(in.mechs.size());
  for (size_t j = 0; j < in.mechs.size(); j++)
    raw[j].assign(rates + j * n_levels, rates + (j + 1) * n_levels);
  in.rate_data = raw;
  // selection relies on non-decreasing cumulative rows
  for (size_t j = 0; j < raw.size(); j++)
    for (int l = 0; l < n_levels; l++)
      if (!(raw[j][l] >= 0.0) || !std::isfinite(raw[j][l]))
        return fail(c, EMCGPU_ERR_ARG,
                    "mechanism %zu, level %d: scatter rates must be finite and non-negative (%g)",
                    j, l, raw[j][l]);
  return finish_tables(c, in, n_levels, max_energy, raw);
}

emcgpu_status emcgpu_build_tables(emcgpu_ctx *c, int instance, int n_levels, double max_energy,
                                  double temperature, const emcgpu_material *mat) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  Instance &in = c->inst[instance];
  if (in.valleys.empty())
    return fail(c, EMCGPU_ERR_STATE, "instance %d: set the model first", instance);
  if (n_levels < 1 || n_levels > EMCGPU_MAX_LEVELS || !(max_energy > 0) || !mat ||
      !(temperature > 0))
    return fail(c, EMCGPU_ERR_ARG,
                "n_levels in [1, %d], max_energy > 0, temperature > 0 and material required",
                EMCGPU_MAX_LEVELS);
  in.temperature = temperature;
  const double dE = max_energy / (double)(std::size_t)n_levels;
  for (size_t j = 0; j < in.mechs.size(); j++) {
    const emcgpu_mech_desc &m = in.mechs[j];
    if (reads_bath(m.kind) && m.bath >= (int)in.baths.size())
      return fail(c, EMCGPU_ERR_STATE, "mechanism %zu reads bath %d: set the baths first", j,
                  m.bath);
  }
  if ((st = pull_baths(c, in)))
    return st;
  std::vector<std::vector<double>> raw(in.mechs.size(), std::vector<double>(n_levels));
  in.rate_data.assign(in.mechs.size(), {});
  in.hot_C.assign(in.mechs.size(), 0.0);
  in.mech_n0.assign(in.mechs.size(), 0.0);
  for (size_t j = 0; j < in.mechs.size(); j++) {
    const emcgpu_mech_desc &m = in.mechs[j];
    const phys::ValleyMass &iv = in.vm[m.valley];
    double C = 0, nq = 0, meff = 0;
    const bool hot =
        m.kind == EMCGPU_MECH_HOT_FROEHLICH_AB || m.kind == EMCGPU_MECH_HOT_FROEHLICH_EM;
    const bool ab = m.kind == EMCGPU_MECH_ZERO_IV_AB || m.kind == EMCGPU_MECH_FIRST_IV_AB ||
                    m.kind == EMCGPU_MECH_FROEHLICH_AB || m.kind == EMCGPU_MECH_HOT_FROEHLICH_AB ||
                    m.kind == EMCGPU_MECH_SCREENED_FROEHLICH_AB ||
                    m.kind == EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB;
    if (is_scr_hot_kind(m.kind) || (in.plasma() && is_scr_kind(m.kind))) {
      // emcScreenedFroehlichInteraction.hpp:128-138 / :184-196 and
      // :257-269 / :333-348: kI and kF kept per level for the rebuilds on
      // the GPU (kI = 0: zero rate); the occupation the Bose one, the bath
      // mean or the bath's window mean
      meff = m.rel_eff_mass * phys::me;
      C = phys::froehlichConst(m.phonon_energy, meff, m.eps_hi, m.eps_lo);
      const bool hotk = is_scr_hot_kind(m.kind);
      const double n0 = hotk ? 0.0 : phys::boseEinstein(m.phonon_energy, temperature);
      const double mean = hotk ? bath_mean(in.baths[m.bath]) : 0.0;
      std::vector<double> data(2 * (size_t)n_levels, 0.0);
      for (int l = 0; l < n_levels; l++) {
        const double E = (double)(std::size_t)(l + 1) * dE;
        double r = 0.0;
        if (ab || E > m.phonon_energy) {
          const double gI = iv.gamma(E);
          const double gF = iv.gamma(ab ? E + m.phonon_energy : E - m.phonon_energy);
          if (gI > 0.0 && gF > 0.0) {
            const double kI = std::sqrt(2.0 * meff * gI * phys::q) / phys::hbar;
            const double kF = std::sqrt(2.0 * meff * gF * phys::q) / phys::hbar;
            data[l] = kI;
            data[n_levels + l] = kF;
            double n = n0;
            if (hotk)
              n = m.q_resolved ? window_nq_host(in.baths[m.bath], in.qs2, mean, std::fabs(kI - kF),
                                                kI + kF)
                               : mean;
            r = (ab ? C * n : C * (n + 1.0)) / kI * phys::screenedLogFactor(kI, kF, in.qs2);
          }
        }
        raw[j][l] = r;
      }
      in.rate_data[j] = std::move(data);
      in.hot_C[j] = C;
      in.mech_n0[j] = n0;
      continue;
    }
    if (hot) {
      // emcHotPhononFroehlichMechanism.hpp:79-91 / :159-173: the Froehlich
      // rate with the bath's mean occupation; kI and ln kept per level for
      // the rebuilds on the GPU (kI = 0: zero rate)
      meff = m.rel_eff_mass * phys::me;
      C = phys::froehlichConst(m.phonon_energy, meff, m.eps_hi, m.eps_lo);
      const double n = bath_mean(in.baths[m.bath]);
      std::vector<double> data(2 * (size_t)n_levels, 0.0);
      for (int l = 0; l < n_levels; l++) {
        const double E = (double)(std::size_t)(l + 1) * dE;
        double r = 0.0;
        if (ab || E > m.phonon_energy) {
          const double gI = iv.gamma(E);
          const double gF = iv.gamma(ab ? E + m.phonon_energy : E - m.phonon_energy);
          if (gI > 0.0 && gF > 0.0) {
            const double kI = std::sqrt(2.0 * meff * gI * phys::q) / phys::hbar;
            const double kF = std::sqrt(2.0 * meff * gF * phys::q) / phys::hbar;
            const double ln = ab ? std::log((kI + kF) / (kF - kI)) : std::log((kI + kF) / (kI - kF));
            data[l] = kI;
            data[n_levels + l] = ln;
            r = ab ? C * n / kI * ln : C * (n + 1.0) / kI * ln;
          }
        }
        raw[j][l] = r;
      }
      in.rate_data[j] = std::move(data);
      in.hot_C[j] = C;
      continue;
    }
    switch (m.kind) {
    case EMCGPU_MECH_ACOUSTIC:
      C = phys::acousticConst(m.coupling, mat->rho, mat->v_sound, temperature);
      break;
    case EMCGPU_MECH_ZERO_IV_AB:
    case EMCGPU_MECH_ZERO_IV_EM:
      C = phys::zeroOrderConst(m.coupling, m.phonon_energy, mat->rho, temperature, m.z, ab);
      break;
    case EMCGPU_MECH_FIRST_IV_AB:
    case EMCGPU_MECH_FIRST_IV_EM:
      C = phys::firstOrderConst(m.coupling, m.phonon_energy, mat->rho, temperature, m.z, ab);
      break;
    default:
      meff = m.rel_eff_mass * phys::me;
      C = phys::froehlichConst(m.phonon_energy, meff, m.eps_hi, m.eps_lo);
      nq = phys::boseEinstein(m.phonon_energy, temperature);
      break;
    }
    for (int l = 0; l < n_levels; l++) {
      const double E = (double)(std::size_t)(l + 1) * dE;
      double r = 0;
      switch (m.kind) {
      case EMCGPU_MECH_ACOUSTIC:
        r = phys::acousticRate(C, iv, E);
        break;
      case EMCGPU_MECH_ZERO_IV_AB:
      case EMCGPU_MECH_ZERO_IV_EM: {
        const phys::ValleyMass &fv = in.vm[m.final_valley];
        r = phys::zeroOrderRate(C, fv, E, m.phonon_energy, fv.bottom - iv.bottom, ab);
        break;
      }
      case EMCGPU_MECH_FIRST_IV_AB:
      case EMCGPU_MECH_FIRST_IV_EM: {
        const phys::ValleyMass &fv = in.vm[m.final_valley];
        r = phys::firstOrderRate(C, iv, fv, E, m.phonon_energy, fv.bottom - iv.bottom, ab);
        break;
      }
      case EMCGPU_MECH_SCREENED_FROEHLICH_AB:
      case EMCGPU_MECH_SCREENED_FROEHLICH_EM:
        r = phys::screenedFroehlichRate(C, nq, meff, iv, E, m.phonon_energy, in.qs2, ab);
        break;
      default:
        r = phys::froehlichRate(C, nq, meff, iv, E, m.phonon_energy, ab);
        break;
      }
      raw[j][l] = r;
    }
    in.rate_data[j] = raw[j];
  }
  return finish_tables(c, in, n_levels, max_energy, raw);
}

// initScatterTables of every instance with the rates evaluated on the GPU
// from the descriptors (one lane per level and mechanism), then normalised
// on the host as emcgpu_build_tables does
emcgpu_status emcgpu_build_tables_all(emcgpu_ctx *c, int n_levels, double max_energy,
                                      const double *temperature, const emcgpu_material *mat) {
  if (!c)
    return EMCGPU_ERR_ARG;
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "emcgpu_build_tables_all needs a device");
  if (n_levels < 1 || n_levels > EMCGPU_MAX_LEVELS || !(max_energy > 0) || !temperature || !mat)
    return fail(c, EMCGPU_ERR_ARG,
                "n_levels in [1, %d], max_energy > 0, temperatures and materials required",
                EMCGPU_MAX_LEVELS);
  const double dE = max_energy / (double)(std::size_t)n_levels;
  std::vector<RateDescJob> jobs;
  std::vector<std::pair<int, int>> owner;  // (instance, mechanism) of each job
  for (int i = 0; i < c->n_inst; i++) {
    Instance &in = c->inst[i];
    if (in.valleys.empty())
      return fail(c, EMCGPU_ERR_STATE, "instance %d: set the model first", i);
    if (!(temperature[i] > 0))
      return fail(c, EMCGPU_ERR_ARG, "instance %d: temperature must be > 0", i);
    bool dynamic = in.plasma();
    for (const emcgpu_mech_desc &m : in.mechs)
      dynamic |= reads_bath(m.kind) || is_scr_hot_kind(m.kind);
    if (dynamic)
      return fail(c, EMCGPU_ERR_ARG,
                  "instance %d: rates that follow a bath or the carriers: use emcgpu_build_tables", i);
    for (size_t j = 0; j < in.mechs.size(); j++) {
      const emcgpu_mech_desc &m = in.mechs[j];
      RateDescJob J{};
      J.kind = m.kind;
      J.absorption = (m.kind == EMCGPU_MECH_ZERO_IV_AB || m.kind == EMCGPU_MECH_FIRST_IV_AB ||
                      m.kind == EMCGPU_MECH_FROEHLICH_AB ||
                      m.kind == EMCGPU_MECH_SCREENED_FROEHLICH_AB) ? 1 : 0;
      J.iv = in.vm[m.valley];
      J.fv = in.vm[(m.kind >= EMCGPU_MECH_ZERO_IV_AB && m.kind <= EMCGPU_MECH_FIRST_IV_EM)
                       ? m.final_valley : m.valley];
      J.hw = m.phonon_energy;
      J.delta = J.fv.bottom - J.iv.bottom;
      J.qs2 = in.qs2;
      const double T = temperature[i];
      switch (m.kind) {
      case EMCGPU_MECH_ACOUSTIC:
        J.C = phys::acousticConst(m.coupling, mat[i].rho, mat[i].v_sound, T);
        break;
      case EMCGPU_MECH_ZERO_IV_AB:
      case EMCGPU_MECH_ZERO_IV_EM:
        J.C = phys::zeroOrderConst(m.coupling, m.phonon_energy, mat[i].rho, T, m.z, J.absorption);
        break;
      case EMCGPU_MECH_FIRST_IV_AB:
      case EMCGPU_MECH_FIRST_IV_EM:
        J.C = phys::firstOrderConst(m.coupling, m.phonon_energy, mat[i].rho, T, m.z, J.absorption);
        break;
      default:
        J.meff = m.rel_eff_mass * phys::me;
        J.C = phys::froehlichConst(m.phonon_energy, J.meff, m.eps_hi, m.eps_lo);
        J.nq = phys::boseEinstein(m.phonon_energy, T);
        break;
      }
      J.out_off = (long long)jobs.size() * n_levels;
      jobs.push_back(J);
      owner.emplace_back(i, (int)j);
    }
  }
  if (jobs.size() > 65535)
    return fail(c, EMCGPU_ERR_ARG, "at most 65535 mechanisms over all instances");
  (void)hipSetDevice(c->device);
  std::vector<double> out(jobs.size() * (size_t)n_levels);
  RateDescJob *d_jobs = nullptr;
  double *d_out = nullptr;
  auto cleanup = [&]() {
    dfreet(d_jobs);
    dfreet(d_out);
  };
  hipError_t e = hipMalloc((void **)&d_jobs, sizeof(RateDescJob) * std::max<size_t>(jobs.size(), 1));
  if (e == hipSuccess)
    e = hipMalloc((void **)&d_out, sizeof(double) * std::max<size_t>(out.size(), 1));
  if (e == hipSuccess && !jobs.empty())
    e = hipMemcpyAsync(d_jobs, jobs.data(), sizeof(RateDescJob) * jobs.size(),
                       hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = launch_rates_desc(d_jobs, (int)jobs.size(), n_levels, dE, d_out, c->stream);
  if (e == hipSuccess && !out.empty())
    e = hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  cleanup();
  if (e != hipSuccess)
    return fail(c, EMCGPU_ERR_HIP, "table construction on the GPU: %s", hipGetErrorString(e));
  std::vector<std::vector<std::vector<double>>> raw(c->n_inst);
  for (int i = 0; i < c->n_inst; i++)
    raw[i].assign(c->inst[i].mechs.size(), {});
  for (size_t k = 0; k < jobs.size(); k++)
    raw[owner[k].first][owner[k].second].assign(out.begin() + (long long)k * n_levels,
                                                  out.begin() + (long long)(k + 1) * n_levels);
  for (int i = 0; i < c->n_inst; i++) {
    Instance &in = c->inst[i];
    in.temperature = temperature[i];
    in.rate_data = raw[i];
    in.hot_C.assign(in.mechs.size(), 0.0);
    in.mech_n0.assign(in.mechs.size(), 0.0);
    emcgpu_status st = finish_tables(c, in, n_levels, max_energy, raw[i]);
    if (st)
      return st;
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_get_tables(const emcgpu_ctx *c, int instance, int valley, double *cum,
                                double *tau0, int *n_tables) {
  emcgpu_status st = check_instance(const_cast<emcgpu_ctx *>(c), instance);
  if (st)
    return st;
  const Instance &in = c->inst[instance];
  if (!in.has_tables)
    return fail(const_cast<emcgpu_ctx *>(c), EMCGPU_ERR_STATE, "no tables");
  if (valley < 0 || valley >= (int)in.valleys.size())
    return fail(const_cast<emcgpu_ctx *>(c), EMCGPU_ERR_ARG, "valley out of range");
  const int nt = (int)in.table_mech[valley].size();
  if (n_tables)
    *n_tables = nt;
  const bool hot_valley = dyn_valley(in, (size_t)valley);
  if (hot_valley && (in.dev_bath || in.dev_scr) && c->device >= 0) {
    // the tables rebuilt on the GPU from the evolving baths
    emcgpu_ctx *cm = const_cast<emcgpu_ctx *>(c);
    (void)hipSetDevice(c->device);
    if (c->params_dirty) {
      const double dt = c->params_dt > 0 ? c->params_dt : 1e-15;
      emcgpu_status s2 = c->precision == EMCGPU_FP32 ? upload_params<float>(cm, dt)
                                                     : upload_params<double>(cm, dt);
      if (s2)
        return s2;
    }
    HIPCHK(cm, hipStreamSynchronize(c->stream));
    std::vector<double> dev((size_t)nt * in.n_levels);
    HIPCHK(cm, hipMemcpy(dev.data(), c->d_cum64 + in.cum_off[valley], sizeof(double) * dev.size(),
                         hipMemcpyDeviceToHost));
    double t0 = 0;
    HIPCHK(cm, hipMemcpy(&t0, c->d_tau0_64 + instance * kMaxV + valley, sizeof(double),
                         hipMemcpyDeviceToHost));
    if (tau0)
      *tau0 = t0;
    if (cum)
      for (int l = 0; l < in.n_levels; l++)
        for (int j = 0; j < nt; j++)
          cum[(size_t)j * in.n_levels + l] = dev[(size_t)l * nt + j];
    return EMCGPU_OK;
  }
  if (tau0)
    *tau0 = in.tau0[valley];
  if (cum)
    std::copy(in.cum[valley].begin(), in.cum[valley].end(), cum);
  return EMCGPU_OK;
}

emcgpu_status emcgpu_set_field(emcgpu_ctx *c, int instance, const double dir[3], double strength,
                               double charge) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!dir)
    return fail(c, EMCGPU_ERR_ARG, "dir must be non-NULL");
  Instance &in = c->inst[instance];
  double d[3] = {dir[0], dir[1], dir[2]};
  // normalize (include/emcUtil.hpp:41-47), then scale twice as the bulk
  // handler does (basicBulkParticleHandler.hpp:100-102, :186)
  double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (n == 0.0)
    return fail(c, EMCGPU_ERR_ARG, "field direction must be non-zero");
  for (int i = 0; i < 3; i++) {
    d[i] /= n;
    in.dir[i] = d[i];
    in.force[i] = d[i] * strength * charge;
  }
  in.has_field = true;
  c->params_dirty = true;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_resize(emcgpu_ctx *c, const size_t *counts) {
  if (!c || !counts)
    return fail(c, EMCGPU_ERR_ARG, "counts must be non-NULL");
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context has no device memory");
  (void)hipSetDevice(c->device);
  // each instance starts on a kUnit boundary (tiles never straddle instances)
  std::vector<long long> start(c->n_inst), cnt(c->n_inst);
  long long total = 0;
  for (int i = 0; i < c->n_inst; i++) {
    start[i] = total;
    cnt[i] = (long long)counts[i];
    total += (cnt[i] + kUnit - 1) / kUnit * kUnit;
  }
  if (total / kTile > (1LL << 30) || total > (long long)UINT32_MAX)
    return fail(c, EMCGPU_ERR_ARG, "too many particles");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->tuned_ppl = c->tuned_spl = 0;
  const size_t rs = real_size(c);
  dfree(c->kx);
  dfree(c->ky);
  dfree(c->kz);
  dfree(c->E);
  dfree(c->tau);
  dfreet(c->vs);
  dfreet(c->d_unit_inst);
  dfreet(c->d_inst_start);
  dfreet(c->d_inst_count);
  c->capacity = total;
  const size_t n = (size_t)std::max<long long>(total, 1);
  HIPCHK(c, hipMalloc(&c->kx, n * rs));
  HIPCHK(c, hipMalloc(&c->ky, n * rs));
  HIPCHK(c, hipMalloc(&c->kz, n * rs));
  HIPCHK(c, hipMalloc(&c->E, n * rs));
  HIPCHK(c, hipMalloc(&c->tau, n * rs));
  HIPCHK(c, hipMalloc((void **)&c->vs, n * sizeof(uint32_t)));
  HIPCHK(c, hipMemsetAsync(c->kx, 0, n * rs, c->stream));
  HIPCHK(c, hipMemsetAsync(c->ky, 0, n * rs, c->stream));
  HIPCHK(c, hipMemsetAsync(c->kz, 0, n * rs, c->stream));
  HIPCHK(c, hipMemsetAsync(c->E, 0, n * rs, c->stream));
  HIPCHK(c, hipMemsetAsync(c->tau, 0, n * rs, c->stream));
  HIPCHK(c, hipMemsetAsync(c->vs, 0, n * sizeof(uint32_t), c->stream));
  if (c->track_pos) {
    emcgpu_status st = alloc_positions(c);
    if (st)
      return st;
  }
  std::vector<int> unit_inst((size_t)std::max<long long>(total / kUnit, 1), 0);
  for (int i = 0; i < c->n_inst; i++)
    for (long long u = start[i] / kUnit; u < (start[i] + cnt[i] + kUnit - 1) / kUnit; u++)
      unit_inst[u] = i;
  HIPCHK(c, hipMalloc((void **)&c->d_unit_inst, unit_inst.size() * sizeof(int)));
  HIPCHK(c, hipMalloc((void **)&c->d_inst_start, c->n_inst * sizeof(long long)));
  HIPCHK(c, hipMalloc((void **)&c->d_inst_count, c->n_inst * sizeof(long long)));
  HIPCHK(c, hipMemcpyAsync(c->d_unit_inst, unit_inst.data(), unit_inst.size() * sizeof(int),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_inst_start, start.data(), c->n_inst * sizeof(long long),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_inst_count, cnt.data(), c->n_inst * sizeof(long long),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->n_inst; i++) {
    c->inst[i].count = (size_t)cnt[i];
    c->inst[i].start = start[i];
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_get_count(const emcgpu_ctx *c, int instance, size_t *count) {
  emcgpu_status st = check_instance(const_cast<emcgpu_ctx *>(c), instance);
  if (st)
    return st;
  if (count)
    *count = c->inst[instance].count;
  return EMCGPU_OK;
}



emcgpu_status emcgpu_set_box(emcgpu_ctx *c, int instance, const double max_pos[3]) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!max_pos || !(max_pos[0] > 0) || !(max_pos[1] > 0) || !(max_pos[2] > 0))
    return fail(c, EMCGPU_ERR_ARG, "the box needs three positive extents");
  if (c->dsim)
    return fail(c, EMCGPU_ERR_STATE, "device simulations keep their own positions");
  for (int d = 0; d < 3; d++)
    c->inst[instance].box[d] = max_pos[d];
  c->params_dirty = true;
  if (c->device >= 0 && !c->track_pos) {
    (void)hipSetDevice(c->device);
    c->track_pos = true;
    c->tuned_ppl = c->tuned_spl = 0;
    if ((st = alloc_positions(c)))
      return st;
  }
  c->track_pos = true;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_set_positions(emcgpu_ctx *c, int instance, const double *pos) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!pos)
    return fail(c, EMCGPU_ERR_ARG, "pos must be non-NULL");
  if (c->device < 0 || !c->track_pos)
    return fail(c, EMCGPU_ERR_STATE, "positions are not tracked (emcgpu_set_box first)");
  (void)hipSetDevice(c->device);
  return c->precision == EMCGPU_FP32 ? positions_t<float>(c, instance, const_cast<double *>(pos), false)
                                     : positions_t<double>(c, instance, const_cast<double *>(pos), false);
}

emcgpu_status emcgpu_get_positions(const emcgpu_ctx *c0, int instance, double *pos) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!pos)
    return fail(c, EMCGPU_ERR_ARG, "pos must be non-NULL");
  if (c->device < 0 || !c->track_pos)
    return fail(c, EMCGPU_ERR_STATE, "positions are not tracked (emcgpu_set_box first)");
  (void)hipSetDevice(c->device);
  return c->precision == EMCGPU_FP32 ? positions_t<float>(c, instance, pos, true)
                                     : positions_t<double>(c, instance, pos, true);
}

emcgpu_status emcgpu_observe(emcgpu_ctx *c, double *out) {
  if (!c || !out)
    return fail(c, EMCGPU_ERR_ARG, "ctx and out must be non-NULL");
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context has no particles");
  if (c->capacity == 0)
    return fail(c, EMCGPU_ERR_STATE, "no particles (call emcgpu_resize first)");
  for (int i = 0; i < c->n_inst; i++)
    if (c->inst[i].count && !c->inst[i].has_tables)
      return fail(c, EMCGPU_ERR_STATE, "instance %d has particles but no scatter tables", i);
  (void)hipSetDevice(c->device);
  return c->precision == EMCGPU_FP32 ? observe_t<float>(c, out) : observe_t<double>(c, out);
}

emcgpu_status emcgpu_set_particles(emcgpu_ctx *c, int instance, const double *k, const double *E,
                                   const double *tau, const int32_t *valley, const int32_t *sub) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!k || !E || !tau || !valley || !sub)
    return fail(c, EMCGPU_ERR_ARG, "all particle arrays are required");
  if (c->inst[instance].valleys.empty())
    return fail(c, EMCGPU_ERR_STATE, "instance %d: set the model first", instance);
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context has no device memory");
  (void)hipSetDevice(c->device);
  return c->precision == EMCGPU_FP32 ? set_particles_t<float>(c, instance, k, E, tau, valley, sub)
                                     : set_particles_t<double>(c, instance, k, E, tau, valley, sub);
}

emcgpu_status emcgpu_get_particles(const emcgpu_ctx *c, int instance, double *k, double *E,
                                   double *tau, int32_t *valley, int32_t *sub) {
  emcgpu_status st = check_instance(const_cast<emcgpu_ctx *>(c), instance);
  if (st)
    return st;
  if (c->device < 0)
    return fail(const_cast<emcgpu_ctx *>(c), EMCGPU_ERR_STATE, "host-only context has no device memory");
  (void)hipSetDevice(c->device);
  return c->precision == EMCGPU_FP32 ? get_particles_t<float>(c, instance, k, E, tau, valley, sub)
                                     : get_particles_t<double>(c, instance, k, E, tau, valley, sub);
}

emcgpu_status emcgpu_init_particles(emcgpu_ctx *c, int instance, uint64_t seed,
                                    double temperature) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!(temperature > 0))
    return fail(c, EMCGPU_ERR_ARG, "temperature must be > 0");
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context has no device memory");
  (void)hipSetDevice(c->device);
  c->inst[instance].temperature = temperature;
  c->params_dirty = true;
  return c->precision == EMCGPU_FP32 ? init_thermal<float>(c, instance, seed)
                                     : init_thermal<double>(c, instance, seed);
}

emcgpu_status emcgpu_init_particles_energy(emcgpu_ctx *c, int instance, uint64_t seed,
                                           double mean_ev, double sigma_ev) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!(mean_ev > 0) || !(sigma_ev >= 0) || !std::isfinite(mean_ev) || !std::isfinite(sigma_ev))
    return fail(c, EMCGPU_ERR_ARG, "mean energy must be > 0 and sigma >= 0");
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context has no device memory");
  (void)hipSetDevice(c->device);
  return c->precision == EMCGPU_FP32 ? init_thermal<float>(c, instance, seed, mean_ev, sigma_ev)
                                     : init_thermal<double>(c, instance, seed, mean_ev, sigma_ev);
}

// emcPlasmonScreening::setQs2 (emcPlasmonScreening.hpp:79): q_s^2 of the
// instance's screened mechanisms, 0 when disabled.  The final-state angle
// reads it from the next step on; the rates at the next table build.
emcgpu_status emcgpu_set_screening(emcgpu_ctx *c, int instance, const emcgpu_screening_desc *d) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!d)
    return fail(c, EMCGPU_ERR_ARG, "screening descriptor required");
  if (d->mode != EMCGPU_SCREENING_FIXED && d->mode != EMCGPU_SCREENING_PLASMA)
    return fail(c, EMCGPU_ERR_ARG, "unknown screening mode %d", d->mode);
  if (!(d->qs2 >= 0.0) || !std::isfinite(d->qs2))
    return fail(c, EMCGPU_ERR_ARG, "q_s^2 must be finite and >= 0 (%g)", d->qs2);
  if (!(d->eps_static > 0.0))
    return fail(c, EMCGPU_ERR_ARG, "eps_static must be > 0");
  if (d->mode == EMCGPU_SCREENING_PLASMA && !(d->volume > 0.0 && std::isfinite(d->volume)))
    return fail(c, EMCGPU_ERR_ARG, "plasma screening needs a volume > 0");
  if (c->device >= 0)
    (void)hipSetDevice(c->device);
  emcgpu_status st2 = pull_screening(c);
  if (st2)
    return st2;
  Instance &in = c->inst[instance];
  const bool was = in.plasma();
  in.scr = *d;
  in.qs2 = d->enabled ? d->qs2 : 0.0;
  // the screened rates keep kI, kF per level in plasma mode: a change of
  // mode needs a new table build
  if (was != in.plasma())
    in.has_tables = false;
  in.dev_scr = false;
  c->params_dirty = true;
  return EMCGPU_OK;
}

// updateScreening + reinitScatterTables before the first step
// (hotCarrierMHP.cpp:630-637): the valley-0 energies summed on the host in
// particle order, as getAvgEnergy does
emcgpu_status emcgpu_update_screening(emcgpu_ctx *c, int instance) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  Instance &in = c->inst[instance];
  if (!in.plasma())
    return fail(c, EMCGPU_ERR_STATE, "instance %d has no plasma screening", instance);
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "no device");
  if (!in.has_tables)
    return fail(c, EMCGPU_ERR_STATE, "instance %d has no scatter tables", instance);
  (void)hipSetDevice(c->device);
  const size_t n = in.count;
  std::vector<double> E(n);
  std::vector<uint32_t> vs(n);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n) {
    const size_t es = c->precision == EMCGPU_FP32 ? sizeof(float) : sizeof(double);
    std::vector<unsigned char> raw(n * es);
    HIPCHK(c, hipMemcpy(raw.data(), (unsigned char *)c->E + (size_t)in.start * es, n * es,
                        hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(vs.data(), c->vs + in.start, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++)
      E[i] = c->precision == EMCGPU_FP32 ? (double)((const float *)raw.data())[i]
                                         : ((const double *)raw.data())[i];
  }
  double sum = 0.0;
  size_t cnt = 0;
  for (size_t i = 0; i < n; i++)
    if ((vs[i] >> 16) == 0) {
      sum += E[i];
      cnt++;
    }
  double qs2 = 0.0;
  if (n != 0) {
    const double ns = (double)n / in.scr.volume;
    const double avg = cnt != 0 ? sum / (double)cnt : 0.0;
    const double ts = 2.0 * avg * phys::q / (3.0 * phys::kB);
    if (ts > 0.0)
      qs2 += ns * phys::q * phys::q / (in.scr.eps_static * phys::eps0 * phys::kB * ts);
  }
  in.qs2 = qs2;
  in.dev_scr = false;
  if ((st = pull_baths(c, in)))
    return st;
  // the next upload lowers the new q_s^2 and rebuilds the tables and the
  // window sums on the GPU from it
  c->params_dirty = true;
  const double dt = c->params_dt > 0 ? c->params_dt : 1e-15;
  return c->precision == EMCGPU_FP32 ? upload_params<float>(c, dt) : upload_params<double>(c, dt);
}

emcgpu_status emcgpu_get_screening(const emcgpu_ctx *c0, int instance, double *qs2) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (!qs2)
    return fail(c, EMCGPU_ERR_ARG, "qs2 pointer required");
  if (c->device >= 0)
    (void)hipSetDevice(c->device);
  if ((st = pull_screening(c)))
    return st;
  *qs2 = c->inst[instance].qs2;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_set_baths(emcgpu_ctx *c, int instance, const emcgpu_bath_desc *baths,
                               int n_baths) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (n_baths < 0 || n_baths > EMCGPU_MAX_BATHS || (n_baths && !baths))
    return fail(c, EMCGPU_ERR_ARG, "n_baths must be in [0, %d]", EMCGPU_MAX_BATHS);
  std::vector<BathHost> hb;
  for (int b = 0; b < n_baths; b++) {
    const emcgpu_bath_desc &d = baths[b];
    if (d.n_bins < 1 || d.n_bins > EMCGPU_MAX_BINS)
      return fail(c, EMCGPU_ERR_ARG, "bath %d: n_bins must be in [1, %d]", b, EMCGPU_MAX_BINS);
    if (!(d.dq > 0) || !(d.tau_lo > 0) || !(d.phonon_energy > 0) || !(d.lattice_temp > 0) ||
        !(d.volume > 0))
      return fail(c, EMCGPU_ERR_ARG,
                  "bath %d: dq, tau_lo, phonon_energy, lattice_temp and volume must be > 0", b);
    if (d.acoustic && (!(d.ac_phonon_energy > 0) || !(d.tau_acoustic > 0)))
      return fail(c, EMCGPU_ERR_ARG, "bath %d: acoustic reservoir needs energy and time > 0", b);
    if (d.acoustic && d.w_ridley > 0 && d.to_phonon_energy > 0 && !(d.tau_to > 0))
      return fail(c, EMCGPU_ERR_ARG, "bath %d: Ridley reservoir needs tau_to > 0", b);
    BathHost B;
    B.d = d;
    B.d.tau_profile = nullptr;
    if (d.tau_profile) {
      B.profile.assign(d.tau_profile, d.tau_profile + d.n_bins);
      for (double t : B.profile)
        if (!(t > 0) || !std::isfinite(t))
          return fail(c, EMCGPU_ERR_ARG, "bath %d: lifetime profile values must be > 0", b);
    }
    // constructor (emcPhononBath.hpp:160-192)
    B.n0 = planck(d.phonon_energy, d.lattice_temp);
    B.nq.assign(d.n_bins, B.n0);
    if (d.acoustic) {
      B.has_ac = true;
      B.n_ac_eq = planck(d.ac_phonon_energy, d.lattice_temp);
      B.n_ac = B.n_ac_eq;
    }
    if (d.acoustic && d.w_ridley > 0 && d.to_phonon_energy > 0) {
      B.has_ridley = true;
      B.w_r = (d.w_ridley > 1.0) ? 1.0 : d.w_ridley;
      B.n_to_eq = planck(d.to_phonon_energy, d.lattice_temp);
      B.n_to = B.n_to_eq;
    }
    hb.push_back(std::move(B));
  }
  Instance &in = c->inst[instance];
  in.baths = std::move(hb);
  in.dev_bath = false;
  if (in.has_hot)
    in.has_tables = false;  // the hot-phonon rates read the new baths
  c->bath_layout_dirty = true;
  c->params_dirty = true;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_get_bath(const emcgpu_ctx *c0, int instance, int bath, double *nq,
                              double *scalars) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  Instance &in = c->inst[instance];
  if (bath < 0 || bath >= (int)in.baths.size())
    return fail(c, EMCGPU_ERR_ARG, "instance %d has no bath %d", instance, bath);
  if (c->device >= 0)
    (void)hipSetDevice(c->device);
  if ((st = pull_baths(c, in)))
    return st;
  const BathHost &B = in.baths[bath];
  if (nq)
    std::copy(B.nq.begin(), B.nq.end(), nq);
  if (scalars) {
    scalars[0] = bath_mean(B);
    scalars[1] = B.n_ac;
    scalars[2] = B.n_to;
    scalars[3] = B.n0;
    // getAcousticTemp / getTOTemp (emcPhononBath.hpp:480-494)
    scalars[4] = (!B.has_ac || B.n_ac <= B.n_ac_eq) ? B.d.lattice_temp
                                                    : planck_temp(B.d.ac_phonon_energy, B.n_ac);
    scalars[5] = (!B.has_ridley || B.n_to <= B.n_to_eq)
                     ? B.d.lattice_temp
                     : planck_temp(B.d.to_phonon_energy, B.n_to);
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_set_bath_state(emcgpu_ctx *c, int instance, int bath, const double *nq,
                                    double n_ac, double n_to) {
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  Instance &in = c->inst[instance];
  if (bath < 0 || bath >= (int)in.baths.size() || !nq)
    return fail(c, EMCGPU_ERR_ARG, "instance %d has no bath %d (or nq is NULL)", instance, bath);
  if (c->device >= 0)
    (void)hipSetDevice(c->device);
  if ((st = pull_baths(c, in)))
    return st;
  BathHost &B = in.baths[bath];
  for (int i = 0; i < B.d.n_bins; i++)
    if (!(nq[i] >= 0) || !std::isfinite(nq[i]))
      return fail(c, EMCGPU_ERR_ARG, "bin %d: occupations must be finite and >= 0", i);
  B.nq.assign(nq, nq + B.d.n_bins);
  B.n_ac = n_ac;
  B.n_to = n_to;
  if (in.dev_bath && c->device >= 0) {
    const BathVar v{B.n_ac, B.n_to, bath_mean(B), 0.0};
    HIPCHK(c, hipMemcpy(c->d_nq + in.bath_off[bath], B.nq.data(), sizeof(double) * B.d.n_bins,
                        hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_bath_var + in.bath_index0 + bath, &v, sizeof(BathVar),
                        hipMemcpyHostToDevice));
  }
  // the tables follow the new state on the next step
  c->params_dirty = true;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_get_bath_tallies(const emcgpu_ctx *c0, int instance, int bath,
                                      uint64_t *n_em, uint64_t *n_abs) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  Instance &in = c->inst[instance];
  if (bath < 0 || bath >= (int)in.baths.size())
    return fail(c, EMCGPU_ERR_ARG, "instance %d has no bath %d", instance, bath);
  const int nb = in.baths[bath].d.n_bins;
  std::vector<unsigned long long> t(2 * (size_t)nb, 0ull), part(2 * (size_t)nb);
  if (in.dev_bath && c->device >= 0) {
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int r = 0; r < kTallyCopies; r++) {
      HIPCHK(c, hipMemcpy(part.data(), c->d_tally + 2 * (r * c->tally_bins + in.bath_off[bath]),
                          sizeof(unsigned long long) * 2 * nb, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < part.size(); i++)
        t[i] += part[i];
    }
  }
  for (int i = 0; i < nb; i++) {
    if (n_em)
      n_em[i] = t[2 * i];
    if (n_abs)
      n_abs[i] = t[2 * i + 1];
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_get_bath_history(const emcgpu_ctx *c0, int instance, int bath,
                                      double *mean_nq, int *n_steps) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  Instance &in = c->inst[instance];
  if (bath < 0 || bath >= (int)in.baths.size())
    return fail(c, EMCGPU_ERR_ARG, "instance %d has no bath %d", instance, bath);
  const int ns = in.dev_bath ? c->bath_hist_steps : 0;
  if (n_steps)
    *n_steps = ns;
  if (mean_nq && ns > 0) {
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(mean_nq, c->d_bath_hist + (in.bath_index0 + bath) * (long long)ns,
                        sizeof(double) * ns, hipMemcpyDeviceToHost));
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_get_bath_energy(const emcgpu_ctx *c0, int instance, int bath,
                                     double *energy, int *n_steps) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  Instance &in = c->inst[instance];
  if (bath < 0 || bath >= (int)in.baths.size())
    return fail(c, EMCGPU_ERR_ARG, "instance %d has no bath %d", instance, bath);
  const int ns = in.dev_bath ? c->bath_hist_steps : 0;
  if (n_steps)
    *n_steps = ns;
  if (energy && ns > 0) {
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double *src = c->d_bath_hist + (size_t)c->n_baths * (size_t)ns +
                        3 * (in.bath_index0 + bath) * (long long)ns;
    HIPCHK(c, hipMemcpy(energy, src, 3 * sizeof(double) * ns, hipMemcpyDeviceToHost));
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_step(emcgpu_ctx *c, const emcgpu_step_args *s) {
  if (!c || !s)
    return fail(c, EMCGPU_ERR_ARG, "ctx and args must be non-NULL");
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context cannot step");
  (void)hipSetDevice(c->device);
  for (int i = 0; i < c->n_inst; i++)
    if (c->inst[i].count && !c->inst[i].has_field)
      return fail(c, EMCGPU_ERR_STATE, "instance %d: set the field first", i);
  return c->precision == EMCGPU_FP32 ? run_steps<float>(c, s) : run_steps<double>(c, s);
}

emcgpu_status emcgpu_reserve(emcgpu_ctx *c, const emcgpu_step_args *s) {
  if (!c || !s)
    return fail(c, EMCGPU_ERR_ARG, "ctx and args must be non-NULL");
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context cannot step");
  (void)hipSetDevice(c->device);
  return c->precision == EMCGPU_FP32 ? run_steps<float>(c, s, nullptr, true)
                                     : run_steps<double>(c, s, nullptr, true);
}

emcgpu_status emcgpu_step_noalloc(emcgpu_ctx *c, const emcgpu_step_args *s, double *obs,
                                  size_t obs_len) {
  if (!c || !s)
    return fail(c, EMCGPU_ERR_ARG, "ctx and args must be non-NULL");
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context cannot step");
  (void)hipSetDevice(c->device);
  for (int i = 0; i < c->n_inst; i++)
    if (c->inst[i].count && !c->inst[i].has_field)
      return fail(c, EMCGPU_ERR_STATE, "instance %d: set the field first", i);
  std::string why;
  double *p = scratch_ptr(c->device, &why);
  if (!p)
    return fail(c, EMCGPU_ERR_HIP, "emcgpu_step_noalloc: no scratch on device %d (%s)",
                c->device, why.c_str());
  const NoAllocBuffers na{p, kNoAllocScratchBytes / sizeof(double), obs, obs_len};
  return c->precision == EMCGPU_FP32 ? run_steps<float>(c, s, &na)
                                     : run_steps<double>(c, s, &na);
}

// Launch-shape autotuner: production trials of the current ensemble for
// each (particles per lane, steps per launch) pair, timed with HIP events on
// the context's stream; the particle state is copied aside once and put back
// before every trial and at the end, so every trial runs the same work and
// the ensemble is left as it was.
emcgpu_status emcgpu_autotune(emcgpu_ctx *c, emcgpu_autotune_args *t) {
  if (!c || !t)
    return fail(c, EMCGPU_ERR_ARG, "ctx and args must be non-NULL");
  if (c->device < 0)
    return fail(c, EMCGPU_ERR_STATE, "host-only context cannot tune");
  if (c->n_baths > 0)
    return fail(c, EMCGPU_ERR_STATE, "autotune is not available with phonon baths");
  if (c->track_pos)
    return fail(c, EMCGPU_ERR_STATE, "autotune is not available with tracked positions "
                                     "(one particle per lane)");
  // plasma screening updates q_s^2 and the screened tables after every
  // step; the trials would leave them at the state of the last trial step
  for (int i = 0; i < c->n_inst; i++)
    if (c->inst[i].plasma())
      return fail(c, EMCGPU_ERR_STATE,
                  "autotune is not available with plasma screening (instance %d)", i);
  if (c->capacity == 0)
    return fail(c, EMCGPU_ERR_STATE, "no particles (call emcgpu_resize first)");
  for (int i = 0; i < c->n_inst; i++)
    if (c->inst[i].count && !c->inst[i].has_field)
      return fail(c, EMCGPU_ERR_STATE, "instance %d: set the field first", i);
  if (t->n_ppl < 0 || t->n_ppl > 4 || t->n_spl < 0 || t->n_spl > 8 || t->repeats < 0 ||
      t->n_steps < 0 || !(t->dt > 0))
    return fail(c, EMCGPU_ERR_ARG, "autotune: dt > 0, n_steps >= 0, repeats >= 0, "
                                   "at most 4 ppl and 8 steps-per-launch candidates");
  std::vector<int> P, S;
  if (t->n_ppl == 0)
    P = {1, 2, 4, 8};
  else
    P.assign(t->ppl, t->ppl + t->n_ppl);
  if (t->n_spl == 0)
    S = {16, 32, 64, 128};
  else
    S.assign(t->spl, t->spl + t->n_spl);
  for (int p : P)
    if (p != 1 && p != 2 && p != 4 && p != 8)
      return fail(c, EMCGPU_ERR_ARG, "autotune: particles per lane must be 1, 2, 4 or 8 (got %d)",
                  p);
  int max_spl = 1;
  for (int k : S) {
    if (k < 1 || k > 4096)
      return fail(c, EMCGPU_ERR_ARG, "autotune: steps per launch must be in [1, 4096] (got %d)", k);
    max_spl = std::max(max_spl, k);
  }
  const int n_steps = t->n_steps > 0 ? t->n_steps : 2 * max_spl;
  const int repeats = t->repeats > 0 ? t->repeats : 1;
  (void)hipSetDevice(c->device);

  // particle state aside: k (3), E, tau in the context precision, vs
  const size_t n = (size_t)c->capacity, rs = real_size(c);
  void *arr[5] = {c->kx, c->ky, c->kz, c->E, c->tau};
  char *snap = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto release = [&]() {
    if (snap)
      (void)hipFree(snap);
    if (e0)
      (void)hipEventDestroy(e0);
    if (e1)
      (void)hipEventDestroy(e1);
  };
  auto copy_state = [&](bool save) -> hipError_t {
    for (int i = 0; i < 5; i++) {
      char *s0 = snap + (size_t)i * n * rs;
      hipError_t e = save ? hipMemcpyAsync(s0, arr[i], n * rs, hipMemcpyDeviceToDevice, c->stream)
                          : hipMemcpyAsync(arr[i], s0, n * rs, hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess)
        return e;
    }
    char *sv = snap + 5 * n * rs;
    return save ? hipMemcpyAsync(sv, c->vs, n * 4, hipMemcpyDeviceToDevice, c->stream)
                : hipMemcpyAsync(c->vs, sv, n * 4, hipMemcpyDeviceToDevice, c->stream);
  };
#define TUNECHK(x)                                                                         do {                                                                                       hipError_t e_ = (x);                                                                     if (e_ != hipSuccess) {                                                                    release();                                                                               return fail(c, EMCGPU_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));            }                                                                                      } while (0)
  TUNECHK(hipMalloc((void **)&snap, n * (5 * rs + 4)));
  TUNECHK(hipEventCreate(&e0));
  TUNECHK(hipEventCreate(&e1));
  TUNECHK(copy_state(true));

  emcgpu_step_args a;
  std::memset(&a, 0, sizeof(a));
  a.dt = t->dt;
  a.n_steps = n_steps;
  a.mode = EMCGPU_MODE_PRODUCTION;
  a.seed = t->seed;
  a.step0 = t->step0;
  t->n_results = 0;
  t->best_ppl = t->best_spl = 0;
  t->best_ms_per_step = 0;
  for (int p : P) {
    for (int k : S) {
      a.particles_per_lane = p;
      a.steps_per_launch = k;
      double best = -1;
      for (int r = 0; r <= repeats; r++) {  // r = 0: untimed warm-up
        TUNECHK(copy_state(false));
        TUNECHK(hipEventRecord(e0, c->stream));
        emcgpu_status st = c->precision == EMCGPU_FP32 ? run_steps<float>(c, &a)
                                                       : run_steps<double>(c, &a);
        if (st) {  // put the ensemble back, keep the trial's error
          const std::string msg = c->err;
          (void)copy_state(false);
          (void)hipStreamSynchronize(c->stream);
          release();
          c->err = msg;
          return st;
        }
        TUNECHK(hipEventRecord(e1, c->stream));
        TUNECHK(hipEventSynchronize(e1));
        float ms = 0;
        TUNECHK(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0 && (best < 0 || ms < best))
          best = ms;
      }
      const double per_step = best / n_steps;
      if (t->n_results < EMCGPU_TUNE_MAX) {
        t->res_ppl[t->n_results] = p;
        t->res_spl[t->n_results] = k;
        t->res_ms_per_step[t->n_results] = per_step;
        t->n_results++;
      }
      if (t->best_ppl == 0 || per_step < t->best_ms_per_step) {
        t->best_ppl = p;
        t->best_spl = k;
        t->best_ms_per_step = per_step;
      }
    }
  }
  TUNECHK(copy_state(false));
  TUNECHK(hipStreamSynchronize(c->stream));
#undef TUNECHK
  release();
  c->tuned_ppl = t->best_ppl;
  c->tuned_spl = t->best_spl;
  c->obs_steps = 0;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_get_observables(const emcgpu_ctx *c0, int instance, double *out, int *n_obs,
                                     int *n_steps) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  emcgpu_status st = check_instance(c, instance);
  if (st)
    return st;
  if (n_obs)
    *n_obs = c->n_obs;
  if (n_steps)
    *n_steps = c->obs_steps;
  if (out && c->obs_steps > 0 && c->device >= 0) {
    (void)hipSetDevice(c->device);
    const size_t m = (size_t)c->obs_steps * c->n_obs;
    HIPCHK(c, hipMemcpyAsync(out, c->d_obs + (size_t)instance * m, m * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_obs_layout(const emcgpu_ctx *c, int *n_obs, int *max_valleys, int *max_deg) {
  if (!c)
    return EMCGPU_ERR_ARG;
  if (n_obs)
    *n_obs = c->n_obs;
  if (max_valleys)
    *max_valleys = c->max_v;
  if (max_deg)
    *max_deg = c->max_deg;
  return EMCGPU_OK;
}

emcgpu_status emcgpu_get_timing(const emcgpu_ctx *c0, double *ms_total, long long *launches,
                                int reset) {
  emcgpu_ctx *c = const_cast<emcgpu_ctx *>(c0);
  if (!c)
    return EMCGPU_ERR_ARG;
  if (ms_total)
    *ms_total = c->ms_total;
  if (launches)
    *launches = c->launches;
  if (reset) {
    c->ms_total = 0;
    c->launches = 0;
  }
  return EMCGPU_OK;
}

emcgpu_status emcgpu_synchronize(emcgpu_ctx *c) {
  if (!c)
    return EMCGPU_ERR_ARG;
  if (c->device < 0)
    return EMCGPU_OK;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EMCGPU_OK;
}

}  // extern "C"

namespace emcgpu {
namespace host {

emcgpu_status upload_model(emcgpu_ctx *c, double dt_s) {
  return c->precision == EMCGPU_FP32 ? upload_params<float>(c, dt_s)
                                     : upload_params<double>(c, dt_s);
}
const void *device_params(const emcgpu_ctx *c) { return c->d_params; }
const void *device_tables(const emcgpu_ctx *c) { return c->d_cum; }

}  // namespace host
}  // namespace emcgpu
