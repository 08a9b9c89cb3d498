// emcgpu_host.h -- host-side state of a context of the C ABI, shared by the
// bulk (emcgpu_capi.cpp) and the device (emcgpu_device_capi.cpp) layers.
#ifndef EMCGPU_HOST_H
#define EMCGPU_HOST_H

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/emcgpu.h"
#include "../include/emcgpu/emcgpuPhysics.hpp"
#include "emcgpu_device.h"

namespace emcgpu {
namespace host {

inline thread_local std::string g_create_error;

struct DeviceSim;  // 2-D device state (emcgpu_device_capi.cpp)
void device_sim_destroy(DeviceSim *d);



// host copy of an LO phonon bath (emcPhononBath<double>, emcPhononBath.hpp:56-192)
struct BathHost {
  emcgpu_bath_desc d{};  // tau_profile pointer not kept
  std::vector<double> profile;
  std::vector<double> nq;
  double n0 = 0, n_ac = 0, n_ac_eq = 0, n_to = 0, n_to_eq = 0, w_r = 0;
  bool has_ac = false, has_ridley = false;
};

struct Instance {
  std::vector<emcgpu_valley_desc> valleys;
  std::vector<emcgpu_mech_desc> mechs;
  std::vector<phys::ValleyMass> vm;
  // normalised tables per valley: [table][level] (FP64 master) and tau0 [s]
  int n_levels = 0;
  double max_energy = 0, dE = 0;
  std::vector<std::vector<double>> cum;     // per valley
  std::vector<std::vector<int>> table_mech;  // per valley: mechanism ids
  std::vector<double> tau0;
  bool has_tables = false;
  double temperature = 300;
  double dir[3] = {1, 0, 0};
  double force[3] = {0, 0, 0};  // N
  bool has_field = false;
  size_t count = 0;
  long long start = 0;
  // phonon baths and the bath-dependent table data: per mechanism its FP64
  // rates per level, or for a hot-phonon mechanism kI[level] then ln[level]
  std::vector<BathHost> baths;
  std::vector<std::vector<double>> rate_data;
  std::vector<double> hot_C;
  std::vector<double> mech_n0;  // screened kinds: Bose occupation of the mode
  bool has_hot = false;
  long long bath_index0 = 0;    // global index of the first bath
  std::vector<long long> bath_off;  // global first bin of each bath
  long long gbin_off = 0;
  int total_bins = 0;
  bool dev_bath = false;        // the device holds the current bath state
  // plasmon screening of the screened Froehlich kinds: q_s^2 [1/m^2]
  emcgpu_screening_desc scr{EMCGPU_SCREENING_FIXED, 0, 1.0, 0.0, 0.0};
  double qs2 = 0;
  bool dev_scr = false;  // the device holds the current q_s^2 (plasma mode)
  bool plasma() const { return scr.mode == EMCGPU_SCREENING_PLASMA && scr.enabled; }
  // periodic bulk box [m] of tracked positions (0: none)
  double box[3] = {0, 0, 0};
  std::vector<long long> cum_off;  // per valley: element offset of the device tables
};

}  // namespace host
}  // namespace emcgpu

struct emcgpu_ctx {
  int device = 0;
  int precision = EMCGPU_FP64;
  int n_inst = 1;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<emcgpu::host::Instance> inst;
  // arena
  long long capacity = 0;  // padded particles (multiple of kUnit)
  void *kx = nullptr, *ky = nullptr, *kz = nullptr, *E = nullptr, *tau = nullptr;
  uint32_t *vs = nullptr;
  // positions of the bulk particles [nm] (allocated once a box is set)
  bool track_pos = false;
  void *px = nullptr, *py = nullptr, *pz = nullptr;
  double *d_obs0 = nullptr;  // ensemble sums of the current state [instance][value]
  int *d_unit_inst = nullptr;
  long long *d_inst_start = nullptr, *d_inst_count = nullptr;
  // parameters and tables
  void *d_params = nullptr;
  void *d_cum = nullptr;
  size_t cum_bytes = 0;
  bool params_dirty = true;
  double params_dt = 0;
  // observables
  int max_v = 1, max_deg = 1, n_obs = 0;
  double *d_partial = nullptr;
  size_t partial_bytes = 0;
  double *d_chunks = nullptr;  // (sum, compensation) pairs of the first reduction pass
  size_t chunk_bytes = 0;
  double *d_obs = nullptr;
  size_t obs_bytes = 0;
  int obs_steps = 0;
  // timing
  double ms_total = 0;
  long long launches = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t gexec = nullptr;  // last captured step sequence
  // launch shape chosen by emcgpu_autotune (0: not tuned)
  int tuned_ppl = 0, tuned_spl = 0;
  // phonon baths of all instances (layout built by upload_baths)
  int n_baths = 0, lds_bins = 0, n_jobs = 0;
  long long tally_bins = 0;  // bins of all baths (one tally copy)
  bool bath_layout_dirty = true;
  emcgpu::BathDev *d_baths = nullptr;
  double *d_nq = nullptr, *d_tau_prof = nullptr;
  unsigned long long *d_tally = nullptr;
  emcgpu::BathVar *d_bath_var = nullptr;
  double *d_bath_hist = nullptr;
  size_t bath_hist_bytes = 0;
  int bath_hist_steps = 0;
  emcgpu::RebuildJob *d_jobs = nullptr;
  emcgpu::BathInstDev *d_groups = nullptr;
  int n_groups = 0;
  double *d_rate_arena = nullptr;
  double *d_rb_max = nullptr;  // [job][kRebuildChunks] of the multi-block rebuild
  double *d_cum64 = nullptr, *d_tau0_64 = nullptr;
  // q_s^2 per instance [1/m^2] and the baths' window sums (FP64 and in the
  // precision of the kinematics): per bath cumW[n_bins + 1], cumWN[n_bins + 1]
  double *d_scr = nullptr;
  double *d_win64 = nullptr;
  void *d_win = nullptr;
  long long win_elems = 0;
  size_t cum64_bytes = 0;
  // 2-D device simulation on instance 0 (null until emcgpu_device_setup)
  emcgpu::host::DeviceSim *dsim = nullptr;
};

#define HIPCHK(c, x)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      return fail((c), EMCGPU_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));      \
  } while (0)

namespace emcgpu {
namespace host {

inline emcgpu_status fail(emcgpu_ctx *c, emcgpu_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c)
    c->err = buf;
  else
    g_create_error = buf;
  return st;
}

#define HIPCHK(c, x)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      return fail((c), EMCGPU_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));      \
  } while (0)

inline size_t real_size(const emcgpu_ctx *c) {
  return c->precision == EMCGPU_FP32 ? sizeof(float) : sizeof(double);
}

inline void dfree(void *&p) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
}
template <typename T>
void dfreet(T *&p) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
}

inline emcgpu_status check_instance(emcgpu_ctx *c, int instance) {
  if (!c)
    return EMCGPU_ERR_ARG;
  if (instance < 0 || instance >= c->n_inst)
    return fail(c, EMCGPU_ERR_ARG, "instance %d out of range [0, %d)", instance, c->n_inst);
  return EMCGPU_OK;
}

// builds (if stale) the parameter blocks and table arena of every instance
// for a step of dt_s seconds, in the context's precision
emcgpu_status upload_model(emcgpu_ctx *c, double dt_s);
const void *device_params(const emcgpu_ctx *c);
const void *device_tables(const emcgpu_ctx *c);

}  // namespace host
}  // namespace emcgpu

#endif
