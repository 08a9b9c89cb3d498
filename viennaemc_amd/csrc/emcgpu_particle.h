// emcgpu_particle.h -- per-particle device functions of the EMC step shared
// by the bulk and the device (2-D, with positions) kernels: precision
// traits, Philox-4x32-10 streams, the reference engine's replay stream,
// valley kinematics in scaled units, final-state samplers and one pass of
// the scatter loop.
#ifndef EMCGPU_PARTICLE_H
#define EMCGPU_PARTICLE_H

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
  static __device__ __forceinline__ double ceil(double x) { return ::ceil(x); }
  static __device__ __forceinline__ double round(double x) { return ::round(x); }
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
  static __device__ __forceinline__ float ceil(float x) { return __builtin_ceilf(x); }
  static __device__ __forceinline__ float round(float x) { return __builtin_roundf(x); }
  // v_sin_f32 / v_cos_f32 take their argument in revolutions
  static __device__ __forceinline__ void sincos2pi(float u, float *s, float *c) {
    *s = __builtin_amdgcn_sinf(u);
    *c = __builtin_amdgcn_cosf(u);
  }
};

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon et al., SC'11)
// ---------------------------------------------------------------------------
// (the 64-bit products compile to one v_mad_u64_u32 each instead of a
// v_mul_lo_u32 + v_mul_hi_u32 pair)
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// a uniform in (0, 1) from the top 24 bits of a word, forced odd: exact in
// FP32, so both precisions see the same value (the FP64 kernels widen it)
__host__ __device__ inline float u01_f32(uint32_t w) {
  return (float)((w >> 8) | 1u) * 0x1p-24f;
}

// the four uniforms of scatter event e: selection, two final-state
// uniforms, free-flight time -- one counter block, one word each, in both
// precisions, so an FP32 and an FP64 run on the same key follow the same
// random path
template <typename R>
struct EventDraws {
  static __device__ __forceinline__ void get(uint32_t e, uint32_t step, uint32_t pid,
                                             uint32_t inst, uint32_t k0, uint32_t k1,
                                             R u[4]) {
    uint32_t c[4] = {e, step, pid, inst};
    philox4x32_10(c, k0, k1);
    u[0] = (R)u01_f32(c[0]);
    u[1] = (R)u01_f32(c[1]);
    u[2] = (R)u01_f32(c[2]);
    u[3] = (R)u01_f32(c[3]);
  }
};

// uniform #j of a (step, particle) stream: word j % 4 of counter block j / 4
template <typename R>
__host__ __device__ inline R stream_uniform(uint32_t j, uint32_t step, uint32_t pid,
                                            uint32_t inst, uint32_t k0, uint32_t k1) {
  uint32_t c[4] = {j >> 2, step, pid, inst};
  philox4x32_10(c, k0, k1);
  return (R)u01_f32(c[j & 3u]);
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
  // sqrt(1 + 4 alpha gamma(E)) = 1 + 2 alpha E for gamma = E (1 + alpha E)
  const R npf = R(1) + R(2) * V.alpha * E;
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

// bin of a phonon wavevector magnitude q (binOf, emcPhononBath.hpp:222-225)
// and its tally in the block's LDS histogram [local bin][emission, absorption]
template <typename R>
__device__ __forceinline__ void hot_record(const InstDev<R> *P, const MechDev<R> &m, R q,
                                           uint32_t *lds_tally) {
  const int b = m.bath1 - 1;
  const int nb = P->bath_nb[b];
  const R r = q / P->bath_dq[b];
  const int bin = (r < (R)nb) ? (int)r : nb - 1;
  atomicAdd(&lds_tally[2 * (P->bath_loff[b] + bin) + (m.shift < R(0) ? 0 : 1)], 1u);
}

// sampleScreenedFroehlichCosTheta (emcScreenedFroehlichInteraction.hpp:
// 65-77): inverse of the cumulative distribution of P(c) ~ 1 / (A' - B c),
// A' = kI^2 + kF^2 + qs2, B = 2 kI kF, with A' - B formed as
// (kI - kF)^2 + qs2 (the same value without the cancellation); isotropic
// 1 - 2u where the form degenerates
template <typename R>
__device__ __forceinline__ R screened_cos_theta(R ki, R kf, R qs2, R u) {
  const R B = R(2) * ki * kf;
  const R dk = ki - kf;
  const R ap = ki * ki + kf * kf + qs2;
  const R num = dk * dk + qs2;
  const R den = ap + B;
  if (!(B > R(0)) || !(num > R(0)) || !(den > R(0)))
    return R(1) - R(2) * u;
  R c = Math<R>::div(ap - den * Math<R>::pow(Math<R>::div(num, den), u), B);
  c = (c < R(1)) ? c : R(1);   // std::min(1, c)
  return (R(-1) < c) ? c : R(-1); // std::max(-1, c)
}

// the staged window sums of the block's instance: after its 2 total_bins
// tally words (8-byte aligned: the dynamic LDS base is 16-byte aligned)
template <typename R>
__device__ __forceinline__ R *lds_window(const InstDev<R> *P, uint32_t *lds_tally) {
  return (R *)(lds_tally + 2 * P->total_bins);  // an even number of words
}

// sampleQ (emcPhononBath.hpp:423-458) over a bath's window sums w (cumW at
// w[0..nb], cumWN at w[nb+1..2nb+1]; their unit cancels): the bins
// [floor(qmin/dq), ceil(qmax/dq)) clamped to the grid, the bin whose prefix
// sum brackets sLo + u (sHi - sLo) found by bisection, then linear
// interpolation inside it; f = N + 1 for emission, N for absorption.  Every
// index stays in [0, nb].
template <typename R>
__device__ __forceinline__ R sample_q(const R *__restrict__ w, int nb, R dq, R qmin, R qmax,
                                      bool emission, R u) {
  if (nb < 2 || !(qmax > qmin))
    return qmin;
  const R flo = Math<R>::floor(Math<R>::div(qmin, dq));
  const R fhi = Math<R>::ceil(Math<R>::div(qmax, dq));
  const int lo = (flo > R(0)) ? ((flo < (R)nb) ? (int)flo : nb) : 0;
  const int hi = (fhi < (R)nb) ? ((fhi > R(0)) ? (int)fhi : 0) : nb;
  if (hi - lo < 1)
    return qmin;
  const R *cw = w, *cwn = w + nb + 1;
  auto S = [&](int i) { return emission ? cwn[i] + cw[i] : cwn[i]; };
  const R slo = S(lo), shi = S(hi);
  const R span = shi - slo;
  if (!(span > R(0)))
    return R(0.5) * (qmin + qmax);
  const R target = slo + u * span;
  int a = lo, b = hi;
  while (b - a > 1) {
    const int mid = (a + b) / 2;
    if (S(mid) <= target)
      a = mid;
    else
      b = mid;
  }
  const R sa = S(a), sb = S(a + 1);
  const R frac = (sb > sa) ? Math<R>::div(target - sa, sb - sa) : R(0.5);
  const R q = ((R)a + frac) * dq;
  const R q1 = (q < qmax) ? q : qmax;  // std::min(qMax, q)
  return (qmin < q1) ? q1 : qmin;      // std::max(qMin, .)
}

// sampleQResolvedCosTheta (emcScreenedFroehlichInteraction.hpp:87-97)
template <typename R>
__device__ __forceinline__ R qres_cos_theta(const R *__restrict__ w, int nb, R dq, R ki, R kf,
                                            bool emission, R u) {
  const R B = R(2) * ki * kf;
  if (!(B > R(0)))
    return R(1) - R(2) * u;
  const R d = ki - kf;
  const R q = sample_q(w, nb, dq, (d < R(0)) ? -d : d, ki + kf, emission, u);
  R c = Math<R>::div(ki * ki + kf * kf - q * q, B);
  c = (c < R(1)) ? c : R(1);
  return (R(-1) < c) ? c : R(-1);
}

// final state of one mechanism: ua/ub are the two final-state uniforms
// (phi / cos-theta for isotropic kinds, exponent / phi for the Froehlich
// kinds)
template <typename R, bool HOT = false, bool SCR = true>
__device__ __forceinline__ void final_state(const InstDev<R> *P, const MechDev<R> &m, R &kx,
                                            R &ky, R &kz, R &E, int &v, int &sub, R ua, R ub,
                                            int zidx, uint32_t *lds_tally = nullptr) {
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
    // emcFroehlichInteraction.hpp:109-128 / :183-200; the hot-phonon kinds
    // (emcHotPhononFroehlichMechanism.hpp:94-117 / :175-199) add the record
    // of |k' - k| in their bath
    const R kx0 = kx, ky0 = ky, kz0 = kz;
    R ct, kn;
    // the screened kinds are compiled into the kernels that may meet them
    bool screened = false;
    if constexpr (SCR)
      screened = m.final_kind == kFinalScreenedFroehlich;
    if (screened) {
      // emcScreenedFroehlichInteraction.hpp:140-152 / :198-211: kI and kF
      // are the valley's |k| before and after the energy change
      const R ki = normk_of(P->valley[v], E);
      E += m.shift;
      kn = normk_of(P->valley[v], E);
      if (m.qres) {
        // the screened hot-phonon kinds with q-resolved angle (:283-290 /
        // :358-365): q through the bath's distribution, its window sums
        // staged in LDS by the step kernels
        const int b = m.bath1 - 1;
        const R *w = (HOT && lds_tally)
                         ? lds_window<R>(P, lds_tally) + 2 * (P->bath_loff[b] + b)
                         : P->bath_win[b];
        ct = qres_cos_theta(w, P->bath_nb[b], P->bath_dq[b], ki, kn, m.shift < R(0), ua);
      } else {
        ct = screened_cos_theta(ki, kn, P->qs2, ua);
      }
    } else {
      const R ei = E;
      E += m.shift;
      const R ef = E;
      const R d = Math<R>::sqrt(ei) - Math<R>::sqrt(ef);
      const R f = Math<R>::div(Math<R>::div(R(2) * Math<R>::sqrt(ei * ef), d), d);
      ct = Math<R>::div(R(1) + f - Math<R>::pow(R(1) + R(2) * f, ua), f);
      ct = (ct < R(1)) ? ct : R(1);   // std::min(1, ct)
      ct = (R(-1) < ct) ? ct : R(-1); // std::max(-1, ct)
      kn = normk_of(P->valley[v], E);
    }
    direction_wrt_k(kx, ky, kz, ct, ub);
    const R kc = Math<R>::sqrt(kx * kx + ky * ky + kz * kz);
    if (kc > R(0)) {
      const R s = Math<R>::div(kn, kc);
      kx = kx * s;
      ky = ky * s;
      kz = kz * s;
    }
    if constexpr (HOT) {
      if (m.bath1 != 0) {
        // subtract(k, kOld) then norm (emcUtil.hpp:25-35, 72-80)
        const R dx = kx - kx0, dy = ky - ky0, dz = kz - kz0;
        hot_record(P, m, Math<R>::sqrt(dx * dx + dy * dy + dz * dz), lds_tally);
      }
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

// free flight of duration t in the current valley/subvalley
// (emcParticleDrift.hpp:11-35: k += F t vogt / hbar in the ellipse frame,
// precomputed here as one device-frame acceleration per subvalley).  With
// POS the position moves too (the leapfrog step of :26-34, in the device
// frame: dx = hbar M (k_old + k_new) / 2 t / m_c(E_new), M = R^T diag(vogt)
// R) and wraps into the periodic box like the bulk handler's driftParticle
// (basicBulkParticleHandler.hpp:600-613); pos[3] in nm.
template <typename R, bool POS = false>
__device__ __forceinline__ void drift(const InstDev<R> *P, R t, R &kx, R &ky, R &kz, R &E, int v,
                                      int sub, R *pos = nullptr) {
  const ValleyDev<R> &V = P->valley[v];
  const R kx0 = kx, ky0 = ky, kz0 = kz;
  kx += V.accel[sub][0] * t;
  ky += V.accel[sub][1] * t;
  kz += V.accel[sub][2] * t;
  E = energy_of(V, kx, ky, kz);
  if constexpr (POS) {
    const R *M = V.hv[sub];
    const R mx = (kx + kx0) * R(0.5), my = (ky + ky0) * R(0.5), mz = (kz + kz0) * R(0.5);
    // cv [m/s per 1/nm] * t [fs] * 1e-6 -> nm
    const R f = Math<R>::div(V.cv * t * R(1e-6), R(1) + R(2) * V.alpha * E);
    R x[3] = {pos[0] + (M[0] * mx + M[1] * my + M[2] * mz) * f,
              pos[1] + (M[1] * mx + M[3] * my + M[4] * mz) * f,
              pos[2] + (M[2] * mx + M[4] * my + M[5] * mz) * f};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const R b = P->box[d];
      pos[d] = (x[d] < R(0)) ? x[d] + b : ((x[d] > b) ? x[d] - b : x[d]);
    }
  }
}

// one scatter event: selection against the tables and the final state
// (emcScatterHandler::scatterParticle, emcScatterHandler.hpp:148-170), then
// the next free-flight time (getNewTau, emcParticleType.hpp:187-189), which
// is returned; e counts the events of the step
template <typename R, int MODE, bool HOT = false, bool SCR = true>
__device__ __forceinline__ R scatter_event(const InstDev<R> *P, const R *__restrict__ cum, R &kx,
                                           R &ky, R &kz, R &E, int &v, int &sub, uint32_t &e,
                                           StepRng<R, MODE> &rng, int &nreal,
                                           uint32_t *lds_tally = nullptr) {
  const int nt = P->n_tables[v];
  R u[4];
  if constexpr (MODE == 0)
    EventDraws<R>::get(e, rng.step, rng.pid, rng.inst, rng.k0, rng.k1, u);
  ++e;
  if (nt > 0) {
    const int level = energy_level(P, E);
    const R *row = cum + P->table_off[v] + (long long)level * nt;
    R usel;
    if constexpr (MODE == 0)
      usel = u[0];
    else
      usel = (R)rng.rs.uniform(0., 1.);
    // emcScatterHandler::scatterParticle (:148-170): self scattering if
    // usel > cum[nt-1], else the first t with cum[t-1] <= usel < cum[t]
    // (the early exit beats a branch-free count of all nt entries: most
    // selections end in the first few, high-rate tables)
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
            // the conditional uniform of the selection draw picks the
            // final subvalley (statistically the reference's rng() % Z)
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
        final_state<R, HOT, SCR>(P, m, kx, ky, kz, E, v, sub, ua, ub, zidx, lds_tally);
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
  return -Math<R>::log(ut) * P->tau0[v];
}

// one pass of the scatter loop of moveParticles: scatter (selection +
// final state), draw tau', drift min(rem, tau'), rem -= tau'
// (basicBulkParticleHandler.hpp:203-212)
template <typename R, int MODE, bool HOT = false, bool SCR = true, bool POS = false>
__device__ __forceinline__ void event_once(const InstDev<R> *P, const R *__restrict__ cum, R &kx,
                                           R &ky, R &kz, R &E, R &tau, R &rem, int &v, int &sub,
                                           uint32_t &e, StepRng<R, MODE> &rng, int &nreal,
                                           uint32_t *lds_tally = nullptr, R *pos = nullptr) {
  const R newTau =
      scatter_event<R, MODE, HOT, SCR>(P, cum, kx, ky, kz, E, v, sub, e, rng, nreal, lds_tally);
  tau += newTau;
  // std::min(rem, newTau)
  drift<R, POS>(P, (newTau < rem) ? newTau : rem, kx, ky, kz, E, v, sub, pos);
  rem -= newTau;
}

// one particle, one step: drift min(tau, dt), then the scatter loop
// (basicBulkParticleHandler.hpp:195-214)
template <typename R, int MODE, bool HOT = false, bool SCR = true, bool POS = false>
__device__ __forceinline__ void move_one(const InstDev<R> *P, const R *__restrict__ cum, R &kx,
                                         R &ky, R &kz, R &E, R &tau, int &v, int &sub,
                                         StepRng<R, MODE> &rng, int &nev, int &nreal,
                                         uint32_t *lds_tally = nullptr, R *pos = nullptr) {
  const R dt = P->dt;
  drift<R, POS>(P, (dt < tau) ? dt : tau, kx, ky, kz, E, v, sub, pos);  // std::min(tau, dt)
  R rem = dt - tau;
  uint32_t e = 0;
  // every pass draws tau' > 0 (the uniform is clamped below 1), so rem
  // strictly decreases; the event cap bounds the loop regardless
  while (rem > R(0) && e < kMaxEventsPerStep)
    event_once<R, MODE, HOT, SCR, POS>(P, cum, kx, ky, kz, E, tau, rem, v, sub, e, rng, nreal,
                                       lds_tally, pos);
  nev += (int)e;
  tau -= dt;
}

}  // namespace emcgpu

#endif
