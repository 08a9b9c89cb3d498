// emcgpu_kernels.hip -- gfx950 kernels of the EMC particle step.
//
// One work-item owns one particle of a 256-particle tile and keeps it in
// registers for all steps fused into a launch, so HBM sees the SoA state
// once per launch.  Each step is the free-flight / scatter loop of
// basicBulkParticleHandler::moveParticles
// (examples/bulkSimulation/basicBulkParticleHandler.hpp:195-221):
//   drift min(tau, dt); while (dt - tau > 0) { scatter; tau += tau';
//   drift min(rem, tau'); rem -= tau' }; tau -= dt
// followed by the per-step ensemble sums (wave ballot / shuffle reductions,
// one FP64 partial per tile per step, Kahan-combined by k_reduce_obs).
//
// Random numbers:
//   production: Philox-4x32-10 keyed by the seed, counter
//               (block, step, particle, instance); one counter block per
//               scatter event in FP32 (4 words -> selection, two final-state
//               uniforms, free-flight time) and two in FP64 (53-bit pairs).
//   replay:     the reference engine's recorded raw 64-bit draws, consumed
//               in the reference's order and mapped like libstdc++'s
//               generate_canonical (FP64 only).
#include <hip/hip_runtime.h>

#include "emcgpu_device.h"

namespace emcgpu {

// ---------------------------------------------------------------------------
// math per precision
// ---------------------------------------------------------------------------
template <typename R>
struct Math;

template <>
struct Math<double> {
  static __device__ __forceinline__ double sqrt(double x) { return __builtin_sqrt(x); }
  static __device__ __forceinline__ double log(double x) { return ::log(x); }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double pow(double a, double b) { return ::pow(a, b); }
  static __device__ __forceinline__ double floor(double x) { return ::floor(x); }
  // sin / cos of 2*pi*u, evaluated as the reference does (phi = 2*pi*u)
  static __device__ __forceinline__ void sincos2pi(double u, double *s, double *c) {
    double phi = 2 * 3.14159265358979323846 * u;
    *s = ::sin(phi);
    *c = ::cos(phi);
  }
};

template <>
struct Math<float> {
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float log(float x) {
    return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
  }
  static __device__ __forceinline__ float div(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  static __device__ __forceinline__ float pow(float a, float b) {
    return __builtin_amdgcn_exp2f(b * __builtin_amdgcn_logf(a));
  }
  static __device__ __forceinline__ float floor(float x) { return __builtin_floorf(x); }
  // v_sin_f32 / v_cos_f32 take their argument in revolutions
  static __device__ __forceinline__ void sincos2pi(float u, float *s, float *c) {
    *s = __builtin_amdgcn_sinf(u);
    *c = __builtin_amdgcn_cosf(u);
  }
};

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon et al., SC'11)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c[0];
    const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2];
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ __forceinline__ float u01_f32(uint32_t w) {
  return (float)((w >> 8) | 1u) * 0x1p-24f;
}
__device__ __forceinline__ double u01_f64(uint32_t hi, uint32_t lo) {
  uint64_t m = ((((uint64_t)hi) << 32) | lo) >> 11;
  return (double)(m | 1ull) * 0x1p-53;
}

// the four uniforms of scatter event e: selection, two final-state
// uniforms, free-flight time
template <typename R>
struct EventDraws;

template <>
struct EventDraws<float> {
  static __device__ __forceinline__ void get(uint32_t e, uint32_t step, uint32_t pid,
                                             uint32_t inst, uint32_t k0, uint32_t k1,
                                             float u[4]) {
    uint32_t c[4] = {e, step, pid, inst};
    philox4x32_10(c, k0, k1);
    u[0] = u01_f32(c[0]);
    u[1] = u01_f32(c[1]);
    u[2] = u01_f32(c[2]);
    u[3] = u01_f32(c[3]);
  }
};

template <>
struct EventDraws<double> {
  static __device__ __forceinline__ void get(uint32_t e, uint32_t step, uint32_t pid,
                                             uint32_t inst, uint32_t k0, uint32_t k1,
                                             double u[4]) {
    uint32_t a[4] = {2u * e, step, pid, inst};
    uint32_t b[4] = {2u * e + 1u, step, pid, inst};
    philox4x32_10(a, k0, k1);
    philox4x32_10(b, k0, k1);
    u[0] = u01_f64(a[0], a[1]);
    u[1] = u01_f64(a[2], a[3]);
    u[2] = u01_f64(b[0], b[1]);
    u[3] = u01_f64(b[2], b[3]);
  }
};

// uniform #j of a (step, particle) stream: FP32 uses one word per uniform,
// FP64 a pair of words
template <typename R>
__device__ __forceinline__ R stream_uniform(uint32_t j, uint32_t step, uint32_t pid,
                                            uint32_t inst, uint32_t k0, uint32_t k1);
template <>
__device__ __forceinline__ float stream_uniform<float>(uint32_t j, uint32_t step, uint32_t pid,
                                                       uint32_t inst, uint32_t k0, uint32_t k1) {
  uint32_t c[4] = {j >> 2, step, pid, inst};
  philox4x32_10(c, k0, k1);
  return u01_f32(c[j & 3u]);
}
template <>
__device__ __forceinline__ double stream_uniform<double>(uint32_t j, uint32_t step, uint32_t pid,
                                                         uint32_t inst, uint32_t k0, uint32_t k1) {
  uint32_t c[4] = {j >> 1, step, pid, inst};
  philox4x32_10(c, k0, k1);
  const int h = (int)(j & 1u) * 2;
  return u01_f64(c[h], c[h + 1]);
}

// recorded raw draws of the reference engine (replay mode)
struct ReplayStream {
  const uint64_t *p;
  long long pos, len;
  bool over;
  __device__ __forceinline__ uint64_t raw() {
    if (pos < len)
      return p[pos++];
    over = true;
    return 0;
  }
  // std::uniform_real_distribution<double>(a, b) over generate_canonical
  __device__ __forceinline__ double uniform(double a, double b) {
#pragma clang fp contract(off)
    double c = (double)raw() * 0x1p-64;
    if (c >= 1.0)
      c = 0x1.fffffffffffffp-1;
    return c * (b - a) + a;
  }
};

// ---------------------------------------------------------------------------
// valley kinematics (scaled units)
// ---------------------------------------------------------------------------
// Kane dispersion: gamma = E(1 + aE) = c2 |k|^2  ->  E = 2 gamma / (1 + sqrt(1 + 4 a gamma))
// (emcNonParabolicAnistropValley.hpp:109-113; exact for a = 0)
template <typename R>
__device__ __forceinline__ R energy_of(const ValleyDev<R> &V, R kx, R ky, R kz) {
  const R g = V.c2 * (kx * kx + ky * ky + kz * kz);
  return Math<R>::div(R(2) * g, R(1) + Math<R>::sqrt(R(1) + R(4) * V.alpha * g));
}

template <typename R>
__device__ __forceinline__ R normk_of(const ValleyDev<R> &V, R E) {
  const R g = E * (R(1) + V.alpha * E);
  return Math<R>::sqrt(Math<R>::div(g, V.c2));
}

template <typename R>
__device__ __forceinline__ R vdir_of(const ValleyDev<R> &V, int sub, R kx, R ky, R kz, R E) {
  const R g = E * (R(1) + V.alpha * E);
  const R npf = Math<R>::sqrt(R(1) + R(4) * V.alpha * g);
  const R d = kx * V.wdir[sub][0] + ky * V.wdir[sub][1] + kz * V.wdir[sub][2];
  return Math<R>::div(V.cv * d, npf);
}

// getEnergyLevel (emcScatterHandler.hpp:237-244): floor(E/dE) - 1 as size_t
template <typename R>
__device__ __forceinline__ int energy_level(const InstDev<R> *P, R E);
template <>
__device__ __forceinline__ int energy_level<double>(const InstDev<double> *P, double E) {
  const double v = ::floor(E / P->dE) - 1.0;
  if (v == -1.0)
    return 0;
  if (!(v >= 0.0) || v > (double)(P->n_levels - 1))
    return P->n_levels - 1;
  return (int)v;
}
template <>
__device__ __forceinline__ int energy_level<float>(const InstDev<float> *P, float E) {
  const float v = __builtin_floorf(E * P->inv_dE) - 1.0f;
  if (v == -1.0f)
    return 0;
  if (!(v >= 0.0f) || v > (float)(P->n_levels - 1))
    return P->n_levels - 1;
  return (int)v;
}

// initRandomDirection (include/emcUtil.hpp:131-139)
template <typename R>
__device__ __forceinline__ void random_direction(R n, R u_phi, R u_cos, R &kx, R &ky, R &kz) {
  R s, c;
  Math<R>::sincos2pi(u_phi, &s, &c);
  const R ct = R(1) - R(2) * u_cos;
  const R st = Math<R>::sqrt(R(1) - ct * ct);
  kx = n * st * c;
  ky = n * st * s;
  kz = n * ct;
}

// initRandomDirectionWithRespectToCurrentK (include/emcUtil.hpp:141-175)
template <typename R>
__device__ __forceinline__ void direction_wrt_k(R &kx, R &ky, R &kz, R cosT, R u_phi) {
  const R kxy = Math<R>::sqrt(kx * kx + ky * ky);
  const R nk = Math<R>::sqrt(kxy * kxy + kz * kz);
  if (nk == R(0)) {
    kx = ky = kz = R(0);
    return;
  }
  const R ct0 = Math<R>::div(kz, nk);
  const R st0 = Math<R>::div(kxy, nk);
  const R cf0 = (kxy > R(0)) ? Math<R>::div(kx, kxy) : R(1);
  const R sf0 = (kxy > R(0)) ? Math<R>::div(ky, kxy) : R(0);
  const R st = Math<R>::sqrt(R(1) - cosT * cosT);
  R s, c;
  Math<R>::sincos2pi(u_phi, &s, &c);
  const R kxp = nk * st * c;
  const R kyp = nk * st * s;
  const R kzp = nk * cosT;
  kx = kxp * cf0 * ct0 - kyp * sf0 + kzp * cf0 * st0;
  ky = kxp * sf0 * ct0 + kyp * cf0 + kzp * sf0 * st0;
  kz = -kxp * st0 + kzp * ct0;
}

// final state of one mechanism: ua/ub are the two final-state uniforms
// (phi / cos-theta for isotropic kinds, exponent / phi for Froehlich)
template <typename R>
__device__ __forceinline__ void final_state(const InstDev<R> *P, const MechDev<R> &m, R &kx,
                                            R &ky, R &kz, R &E, int &v, int &sub, R ua, R ub,
                                            int zidx) {
  if (m.final_kind == kFinalIsotropicElastic) {
    // emcAcousticScatterMechanism.hpp:70-72
    const R n = Math<R>::sqrt(kx * kx + ky * ky + kz * kz);
    random_direction(n, ua, ub, kx, ky, kz);
  } else if (m.final_kind == kFinalIsotropicInelastic) {
    // emcZeroOrderInterValleyScatterMechanism.hpp:119-129 / :262-272
    sub = m.final_sub[sub][zidx];
    v = m.final_valley;
    E += m.shift;
    const R n = normk_of(P->valley[v], E);
    random_direction(n, ua, ub, kx, ky, kz);
  } else {
    // emcFroehlichInteraction.hpp:109-128 / :183-200
    const R ei = E;
    E += m.shift;
    const R ef = E;
    const R d = Math<R>::sqrt(ei) - Math<R>::sqrt(ef);
    const R f = Math<R>::div(Math<R>::div(R(2) * Math<R>::sqrt(ei * ef), d), d);
    R ct = Math<R>::div(R(1) + f - Math<R>::pow(R(1) + R(2) * f, ua), f);
    ct = (ct < R(1)) ? ct : R(1);   // std::min(1, ct)
    ct = (R(-1) < ct) ? ct : R(-1); // std::max(-1, ct)
    const R kn = normk_of(P->valley[v], E);
    direction_wrt_k(kx, ky, kz, ct, ub);
    const R kc = Math<R>::sqrt(kx * kx + ky * ky + kz * kz);
    if (kc > R(0)) {
      const R s = Math<R>::div(kn, kc);
      kx = kx * s;
      ky = ky * s;
      kz = kz * s;
    }
  }
}

// ---------------------------------------------------------------------------
// one particle, one step
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxEventsPerStep = 1u << 16;

// largest value below 1 of each precision
template <typename R>
struct UnitBelow;
template <>
struct UnitBelow<float> {
  static constexpr float value = 0x1.fffffep-1f;
};
template <>
struct UnitBelow<double> {
  static constexpr double value = 0x1.fffffffffffffp-1;
};

template <typename R, int MODE>
struct StepRng;

template <typename R>
struct StepRng<R, 0> {  // production
  uint32_t step, pid, inst, k0, k1;
};
template <typename R>
struct StepRng<R, 1> {  // replay
  ReplayStream rs;
};

template <typename R, int MODE>
__device__ __forceinline__ void move_one(const InstDev<R> *P, const R *__restrict__ cum, R &kx,
                                         R &ky, R &kz, R &E, R &tau, int &v, int &sub,
                                         StepRng<R, MODE> &rng, int &nev, int &nreal) {
  const R dt = P->dt;
  {
    const ValleyDev<R> &V = P->valley[v];
    const R t = (dt < tau) ? dt : tau;  // std::min(tau, dt)
    kx += V.accel[sub][0] * t;
    ky += V.accel[sub][1] * t;
    kz += V.accel[sub][2] * t;
    E = energy_of(V, kx, ky, kz);
  }
  R rem = dt - tau;
  uint32_t e = 0;
  // every pass draws tau' > 0 (the uniform is clamped below 1), so rem
  // strictly decreases; the event cap bounds the loop regardless
  while (rem > R(0) && e < kMaxEventsPerStep) {
    const int nt = P->n_tables[v];
    R u[4];
    bool have = false;
    if constexpr (MODE == 0) {
      EventDraws<R>::get(e, rng.step, rng.pid, rng.inst, rng.k0, rng.k1, u);
      have = true;
    }
    ++e;
    ++nev;
    if (nt > 0) {
      const int level = energy_level(P, E);
      const R *row = cum + P->table_off[v] + (long long)level * nt;
      R usel;
      if constexpr (MODE == 0)
        usel = u[0];
      else
        usel = (R)rng.rs.uniform(0., 1.);
      if (!(usel > row[nt - 1])) {
        R lower = R(0), upper = R(0);
        int j = -1;
        for (int t = 0; t < nt; ++t) {
          upper = row[t];
          if (usel >= lower && usel < upper) {
            j = t;
            break;
          }
          lower = upper;
        }
        if (j >= 0) {
          const MechDev<R> &m = P->mech[v][j];
          R ua, ub;
          int zidx = 0;
          if constexpr (MODE == 0) {
            ua = u[1];
            ub = u[2];
            if (m.final_kind == kFinalIsotropicInelastic) {
              const R uc = Math<R>::div(usel - lower, upper - lower);
              zidx = (int)(uc * (R)m.z);
              zidx = zidx < m.z ? zidx : m.z - 1;
            }
          } else {
            // reference draw order (g++ evaluates call arguments right to
            // left): IV raw subvalley draw, then cos(theta), then phi;
            // Froehlich: exponent, then phi
            if (m.final_kind == kFinalIsotropicElastic) {
              ub = (R)rng.rs.uniform(0., 1.);
              ua = (R)rng.rs.uniform(0., 1.);
            } else if (m.final_kind == kFinalIsotropicInelastic) {
              zidx = (int)(rng.rs.raw() % (uint64_t)m.z);
              ub = (R)rng.rs.uniform(0., 1.);
              ua = (R)rng.rs.uniform(0., 1.);
            } else {
              ua = (R)rng.rs.uniform(0., 1.);
              ub = (R)rng.rs.uniform(0., 1.);
            }
          }
          final_state(P, m, kx, ky, kz, E, v, sub, ua, ub, zidx);
          ++nreal;
        }
      }
    }
    // getNewTau (emcParticleType.hpp:187-189): -ln U[1e-6, 1) * tau0
    R ut;
    if constexpr (MODE == 0) {
      ut = u[3] * R(1 - 1e-6) + R(1e-6);
      ut = (ut < UnitBelow<R>::value) ? ut : UnitBelow<R>::value;
    } else {
      ut = (R)rng.rs.uniform(1e-6, 1.);
    }
    (void)have;
    const R newTau = -Math<R>::log(ut) * P->tau0[v];
    tau += newTau;
    const ValleyDev<R> &V = P->valley[v];
    const R t = (newTau < rem) ? newTau : rem;  // std::min(rem, newTau)
    kx += V.accel[sub][0] * t;
    ky += V.accel[sub][1] * t;
    kz += V.accel[sub][2] * t;
    E = energy_of(V, kx, ky, kz);
    rem -= newTau;
  }
  tau -= dt;
}

// ---------------------------------------------------------------------------
// wave-level reductions (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    x += __shfl_xor(x, off, 64);
  return x;
}
__device__ __forceinline__ int wave_sum_i(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    x += __shfl_xor(x, off, 64);
  return x;
}

// ---------------------------------------------------------------------------
// the step kernel
// ---------------------------------------------------------------------------
template <typename R, int MODE>
__global__ __launch_bounds__(kTile) void k_bulk_steps(StepArgs<R> a) {
  __shared__ double wsum[kTile / 64][2 + kMaxV * (3 + kMaxDeg)];
  const int tile = blockIdx.x;
  const int inst = a.tile_inst[tile];
  const InstDev<R> *P = a.inst + inst;
  const long long base = a.inst_start[inst];
  const long long local = (long long)tile * kTile + threadIdx.x - base;
  const bool valid = local < a.inst_count[inst];
  const long long gi = base + local;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;

  R kx = 0, ky = 0, kz = 0, E = 0, tau = 0;
  int v = 0, sub = 0;
  if (valid) {
    kx = a.kx[gi];
    ky = a.ky[gi];
    kz = a.kz[gi];
    E = a.E[gi];
    tau = a.tau[gi];
    const uint32_t vs = a.vs[gi];
    v = (int)(vs >> 16);
    sub = (int)(vs & 0xffffu);
  }
  StepRng<R, MODE> rng;
  double grain = 1.0;
  if constexpr (MODE == 0) {
    rng.pid = (uint32_t)local;
    rng.inst = (uint32_t)inst;
    rng.k0 = a.key0;
    rng.k1 = a.key1;
  } else {
    rng.rs.p = a.draws + (valid ? local : 0) * a.draw_stride;
    rng.rs.len = a.draw_stride;
    rng.rs.pos = valid ? (long long)a.cursor[local] : 0;
    rng.rs.over = false;
    if (valid && a.grain_tau)
      grain = a.grain_tau[local];
  }

  const int nv = P->n_valleys;
  for (int s = 0; s < a.n_steps; ++s) {
    int nev = 0, nreal = 0;
    if constexpr (MODE == 0)
      rng.step = a.step0 + (uint32_t)s;
    if (valid) {
      move_one<R, MODE>(P, a.cum, kx, ky, kz, E, tau, v, sub, rng, nev, nreal);
      if constexpr (MODE == 1) {
        // grain bookkeeping (basicBulkParticleHandler.hpp:213-220), no
        // grain mechanism: a new grainTau of -ln U[1e-6,1) * 1 s
        grain -= a.dt_s;
        if (grain <= 0)
          grain = -::log(rng.rs.uniform(1e-6, 1.)) * 1.0;
      }
      if (a.tr_valley && inst == 0) {
        const long long o = (long long)s * a.tr_n + local;
        a.tr_valley[o] = v;
        a.tr_sub[o] = sub;
        a.tr_events[o] = nev;
        a.tr_k[3 * o] = (double)kx;
        a.tr_k[3 * o + 1] = (double)ky;
        a.tr_k[3 * o + 2] = (double)kz;
        a.tr_E[o] = (double)E;
      }
    }
    // ---- per-step ensemble sums of this tile ----
    const int ev = wave_sum_i(nev);
    const int re = wave_sum_i(nreal);
    double *ws = wsum[wave];
    if (lane == 0) {
      ws[0] = (double)ev;
      ws[1] = (double)re;
    }
    for (int vv = 0; vv < a.max_v; ++vv) {
      const bool in = valid && v == vv && vv < nv;
      double se = 0.0, sv = 0.0;
      if (vv < nv) {
        const double de = in ? (double)E : 0.0;
        const double dv = in ? (double)vdir_of(P->valley[vv], sub, kx, ky, kz, E) : 0.0;
        se = wave_sum(de);
        sv = wave_sum(dv);
      }
      const unsigned long long bal = __ballot(in);
      if (lane == 0) {
        ws[2 + 3 * vv] = (double)__popcll(bal);
        ws[3 + 3 * vv] = se;
        ws[4 + 3 * vv] = sv;
      }
      for (int ss = 0; ss < a.max_deg; ++ss) {
        const unsigned long long b2 = __ballot(in && sub == ss);
        if (lane == 0)
          ws[2 + 3 * a.max_v + vv * a.max_deg + ss] = (double)__popcll(b2);
      }
    }
    __syncthreads();
    if (threadIdx.x < a.n_obs) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kTile / 64; ++w)
        t += wsum[w][threadIdx.x];
      a.partial[((long long)s * a.n_tiles + tile) * a.n_obs + threadIdx.x] = t;
    }
    __syncthreads();
  }

  if (valid) {
    a.kx[gi] = kx;
    a.ky[gi] = ky;
    a.kz[gi] = kz;
    a.E[gi] = E;
    a.tau[gi] = tau;
    a.vs[gi] = ((uint32_t)v << 16) | (uint32_t)sub;
    if constexpr (MODE == 1) {
      a.cursor[local] = (uint64_t)rng.rs.pos;
      if (a.grain_tau)
        a.grain_tau[local] = grain;
      if (rng.rs.over)
        atomicAdd(a.overflow, 1);
    }
  }
}

// ---------------------------------------------------------------------------
// per-(instance, step) Kahan / TwoSum reduction of the tile partials
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce_obs(const double *__restrict__ partial,
                                                    int n_steps, int n_tiles, int n_obs,
                                                    const int *__restrict__ inst_tile0,
                                                    const int *__restrict__ inst_ntiles,
                                                    int n_inst, double *__restrict__ out,
                                                    int out_step0, int out_total) {
  __shared__ double ss[256], sc[256];
  const int s = blockIdx.x / n_inst;
  const int inst = blockIdx.x % n_inst;
  const int t0 = inst_tile0[inst], nt = inst_ntiles[inst];
  for (int o = 0; o < n_obs; ++o) {
    double sum = 0.0, comp = 0.0;
    for (int t = threadIdx.x; t < nt; t += 256) {
      const double x = partial[((long long)s * n_tiles + t0 + t) * n_obs + o];
      // Kahan-Babuska (Neumaier) step
      const double y = sum + x;
      if (fabs(sum) >= fabs(x))
        comp += (sum - y) + x;
      else
        comp += (x - y) + sum;
      sum = y;
    }
    ss[threadIdx.x] = sum;
    sc[threadIdx.x] = comp;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        const double a = ss[threadIdx.x], b = ss[threadIdx.x + w];
        const double y = a + b;
        const double bb = y - a;
        const double err = (a - (y - bb)) + (b - bb);  // TwoSum
        ss[threadIdx.x] = y;
        sc[threadIdx.x] += sc[threadIdx.x + w] + err;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0)
      out[((long long)inst * out_total + out_step0 + s) * n_obs + o] = ss[0] + sc[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// thermal initialisation (emcElectron.hpp:75-90, emcParticleInitialization.hpp:36-51)
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kTile) void k_init_thermal(StepArgs<R> a) {
  const int tile = a.tile_base + blockIdx.x;
  const int inst = a.tile_inst[tile];
  const InstDev<R> *P = a.inst + inst;
  const long long base = a.inst_start[inst];
  const long long local = (long long)tile * kTile + threadIdx.x - base;
  if (local >= a.inst_count[inst])
    return;
  const long long gi = base + local;
  const uint32_t st = 0xFFFFFFFFu, pid = (uint32_t)local, in = (uint32_t)inst;
  const R lo = R(1e-6), sc = R(1 - 1e-6);
  const R uv = stream_uniform<R>(0, st, pid, in, a.key0, a.key1) * sc + lo;
  const R us = stream_uniform<R>(1, st, pid, in, a.key0, a.key1) * sc + lo;
  const R ue = stream_uniform<R>(2, st, pid, in, a.key0, a.key1) * sc + lo;
  const R uphi = stream_uniform<R>(3, st, pid, in, a.key0, a.key1);
  const R ucos = stream_uniform<R>(4, st, pid, in, a.key0, a.key1);
  const R ut = stream_uniform<R>(5, st, pid, in, a.key0, a.key1) * sc + lo;
  int v = (int)Math<R>::floor((R)P->n_valleys * uv);
  v = v < P->n_valleys ? v : P->n_valleys - 1;
  const ValleyDev<R> &V = P->valley[v];
  int sub = (int)Math<R>::floor((R)V.deg * us);
  sub = sub < V.deg ? sub : V.deg - 1;
  const R E = R(-1.5) * P->vt * Math<R>::log(ue);
  R kx, ky, kz;
  random_direction(normk_of(V, E), uphi, ucos, kx, ky, kz);
  a.kx[gi] = kx;
  a.ky[gi] = ky;
  a.kz[gi] = kz;
  a.E[gi] = E;
  a.tau[gi] = -Math<R>::log(ut) * P->tau0[v];
  a.vs[gi] = ((uint32_t)v << 16) | (uint32_t)sub;
}

// ---------------------------------------------------------------------------
// launch wrappers (called by the C ABI layer)
// ---------------------------------------------------------------------------
template <typename R>
hipError_t launch_steps(const StepArgs<R> &a, int mode, hipStream_t st) {
  if (a.n_tiles <= 0)
    return hipSuccess;
  if (mode == 0)
    hipLaunchKernelGGL((k_bulk_steps<R, 0>), dim3(a.n_tiles), dim3(kTile), 0, st, a);
  else
    hipLaunchKernelGGL((k_bulk_steps<R, 1>), dim3(a.n_tiles), dim3(kTile), 0, st, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_init(const StepArgs<R> &a, hipStream_t st) {
  if (a.n_tiles <= 0)
    return hipSuccess;
  hipLaunchKernelGGL((k_init_thermal<R>), dim3(a.n_tiles), dim3(kTile), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_reduce(const double *partial, int n_steps, int n_tiles, int n_obs,
                         const int *inst_tile0, const int *inst_ntiles, int n_inst, double *out,
                         int out_step0, int out_total, hipStream_t st) {
  if (n_steps <= 0 || n_inst <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_reduce_obs, dim3(n_steps * n_inst), dim3(256), 0, st, partial, n_steps,
                     n_tiles, n_obs, inst_tile0, inst_ntiles, n_inst, out, out_step0, out_total);
  return hipGetLastError();
}

template hipError_t launch_steps<float>(const StepArgs<float> &, int, hipStream_t);
template hipError_t launch_steps<double>(const StepArgs<double> &, int, hipStream_t);
template hipError_t launch_init<float>(const StepArgs<float> &, hipStream_t);
template hipError_t launch_init<double>(const StepArgs<double> &, hipStream_t);

}  // namespace emcgpu
