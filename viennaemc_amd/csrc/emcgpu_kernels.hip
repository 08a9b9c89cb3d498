// emcgpu_kernels.hip -- gfx950 kernels of the EMC particle step.
//
// A work-item owns PPT particles of a tile of 256 * PPT and keeps them in
// registers for all steps fused into a launch, so HBM sees the SoA state
// once per launch.  Each step is the free-flight / scatter loop of
// basicBulkParticleHandler::moveParticles
// (examples/bulkSimulation/basicBulkParticleHandler.hpp:195-221):
//   drift min(tau, dt); while (dt - tau > 0) { scatter; tau += tau';
//   drift min(rem, tau'); rem -= tau' }; tau -= dt
// followed by the per-step ensemble sums (wave ballot / shuffle reductions,
// one FP64 partial per tile per step, combined by the compensated two-pass
// reduction k_reduce_chunks / k_reduce_final).
//
// Random numbers:
//   production: Philox-4x32-10 keyed by the seed, counter
//               (block, step, particle, instance); one counter block per
//               scatter event (4 words -> selection, two final-state
//               uniforms, free-flight time), 24-bit uniforms that FP64
//               widens, so both precisions follow the same random path.
//   replay:     the reference engine's recorded raw 64-bit draws, consumed
//               in the reference's order and mapped like libstdc++'s
//               generate_canonical (FP64 only).
#include <hip/hip_runtime.h>

#include "emcgpu_particle.h"

namespace emcgpu {

// ---------------------------------------------------------------------------
// wave-level reductions (wave64)
// ---------------------------------------------------------------------------
// DPP moves (GCN/CDNA data-parallel primitives): quad_perm [1,0,3,2] and
// [2,3,0,1], row_shr:4 and :8 within rows of 16 lanes, row_bcast:15 (rows 1
// and 3) and row_bcast:31 (rows 2 and 3); lanes without a source read 0, so
// after the six steps lane 63 holds the wave's sum, which readlane makes
// uniform.  VALU moves instead of LDS permutes.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = dpp_i<CTRL, ROWS>((int)(b & 0xffffffffll));
  const int hi = dpp_i<CTRL, ROWS>((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double wave_sum(double x) {
  x = x + dpp_d<0xb1, 0xf>(x);
  x = x + dpp_d<0x4e, 0xf>(x);
  x = x + dpp_d<0x114, 0xf>(x);
  x = x + dpp_d<0x118, 0xf>(x);
  x = x + dpp_d<0x142, 0xa>(x);
  x = x + dpp_d<0x143, 0xc>(x);
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ int wave_sum_i(int x) {
  x = x + dpp_i<0xb1, 0xf>(x);
  x = x + dpp_i<0x4e, 0xf>(x);
  x = x + dpp_i<0x114, 0xf>(x);
  x = x + dpp_i<0x118, 0xf>(x);
  x = x + dpp_i<0x142, 0xa>(x);
  x = x + dpp_i<0x143, 0xc>(x);
  return __builtin_amdgcn_readlane(x, 63);
}

constexpr int kObsMax = 2 + kMaxV * (3 + kMaxDeg);
static_assert(kObsMax <= 64, "one lane per observable of a wave");

// Per-step ensemble sums of one wave (each lane holds PPT particles): event
// counts, per-valley count / sum E / sum v.dir, per-(valley, subvalley)
// counts.  Lanes sum their own particles in FP64, the wave reduces with
// shuffles and ballots, and lane o writes value o of the wave's partial:
// partial[step][obs][tile * kSubs + wave].  No block barrier: the waves of a
// block run their steps independently.
template <typename R, int PPT>
__device__ __forceinline__ void tile_observables(const StepArgs<R> &a, const InstDev<R> *P,
                                                 int tile, int s, const bool (&valid)[PPT],
                                                 const int (&v)[PPT], const int (&sub)[PPT],
                                                 const R (&kx)[PPT], const R (&ky)[PPT],
                                                 const R (&kz)[PPT], const R (&E)[PPT], int nev,
                                                 int nreal) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nv = P->n_valleys;
  double mine = 0.0;  // the value this lane writes
  const int ev = wave_sum_i(nev);
  const int re = wave_sum_i(nreal);
  mine = (lane == 0) ? (double)ev : mine;
  mine = (lane == 1) ? (double)re : mine;
  for (int vv = 0; vv < a.max_v; ++vv) {
    double se = 0.0, sv = 0.0;
    int cnt = 0;
    if (vv < nv) {
      double de = 0.0, dv = 0.0;
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const bool in = valid[j] && v[j] == vv;
        de += in ? (double)E[j] : 0.0;
        dv += in ? (double)vdir_of(P->valley[vv], sub[j], kx[j], ky[j], kz[j], E[j]) : 0.0;
        cnt += (int)__popcll(__ballot(in));
      }
      se = wave_sum(de);
      sv = wave_sum(dv);
    }
    mine = (lane == 2 + 3 * vv) ? (double)cnt : mine;
    mine = (lane == 3 + 3 * vv) ? se : mine;
    mine = (lane == 4 + 3 * vv) ? sv : mine;
    for (int ss = 0; ss < a.max_deg; ++ss) {
      int c2 = 0;
#pragma unroll
      for (int j = 0; j < PPT; ++j)
        c2 += (int)__popcll(__ballot(valid[j] && v[j] == vv && sub[j] == ss));
      mine = (lane == 2 + 3 * a.max_v + vv * a.max_deg + ss) ? (double)c2 : mine;
    }
  }
  if (lane < a.n_obs)
    a.partial[((long long)s * a.n_obs + lane) * ((long long)a.n_tiles * kSubs) +
              (long long)tile * kSubs + wave] = mine;
}

// instance of a kernel tile of tile_size particles (tiles never straddle
// instances: instances start on kUnit boundaries and kUnit % tile_size == 0)
__device__ __forceinline__ int tile_instance(const int *unit_inst, long long tile, int tile_size) {
  return unit_inst[(tile * tile_size) / kUnit];
}

// ---------------------------------------------------------------------------
// phonon-bath tallies: each block counts its events in an LDS histogram of
// its instance's bins and adds the non-zero counts to the global 64-bit
// tallies once at the end
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void tally_clear(const InstDev<R> *P, uint32_t *lds_tally) {
  for (int i = threadIdx.x; i < 2 * P->total_bins; i += kTile)
    lds_tally[i] = 0u;
  // the baths' window sums, read by the bisection of the q-resolved angle,
  // from LDS: they change only between steps (one step per launch here)
  if (P->win_lds) {
    R *dst = lds_window<R>(P, lds_tally);
    const R *src = P->bath_win[0];
    for (int i = threadIdx.x; i < 2 * (P->total_bins + P->n_baths); i += kTile)
      dst[i] = src[i];
  }
  __syncthreads();
}

template <typename R>
__device__ __forceinline__ void tally_flush(const InstDev<R> *P, const uint32_t *lds_tally,
                                            unsigned long long *tally, long long tally_bins) {
  __syncthreads();
  unsigned long long *g =
      tally + 2 * ((long long)(blockIdx.x % kTallyCopies) * tally_bins + P->gbin_off);
  for (int i = threadIdx.x; i < 2 * P->total_bins; i += kTile) {
    const uint32_t c = lds_tally[i];
    if (c)
      atomicAdd(&g[i], (unsigned long long)c);
  }
}

// ---------------------------------------------------------------------------
// production step kernel: PPT particles per lane
//
// Each lane owns PPT particles of a tile of 256*PPT (particle j of lane t is
// tile_first + j*256 + t, so every load and store is a contiguous 256-wide
// access).  Per step all PPT particles drift min(tau, dt); then the lane
// works through its own pending scatter events one at a time, pulling the
// state of the next particle that still has time left into working
// registers.  A wavefront therefore runs the scatter body while any lane
// has any pending event among its PPT particles, which keeps most lanes
// busy in the body instead of ~20 % (one particle per lane at ~0.24 events
// per step).  No barriers are involved in the event loop.
// ---------------------------------------------------------------------------
// the block's instance parameters copied into LDS: every per-lane indexed
// read of the step (subvalley accelerations, table offsets, mechanism
// descriptors, velocity weights) then takes a 32-bit LDS address instead of
// a 64-bit global address computed on the vector ALU
template <typename R>
__device__ __forceinline__ void stage_params(InstDev<R> &dst, const InstDev<R> *src) {
  static_assert(sizeof(InstDev<R>) % 8 == 0, "8-byte copy");
  const unsigned long long *s = reinterpret_cast<const unsigned long long *>(src);
  unsigned long long *d = reinterpret_cast<unsigned long long *>(&dst);
  for (int i = threadIdx.x; i < (int)(sizeof(InstDev<R>) / 8); i += kTile)
    d[i] = s[i];
  __syncthreads();
}

template <typename R, int PPT, bool HOT, bool POS>
__global__ __launch_bounds__(kTile) void k_bulk_multi(StepArgs<R> a) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds_tally[];
  __shared__ InstDev<R> lds_params;
  constexpr int TS = kTile * PPT;
  const int tile = blockIdx.x;
  const int inst = tile_instance(a.unit_inst, tile, TS);
  stage_params(lds_params, a.inst + inst);
  const InstDev<R> *P = &lds_params;
  if constexpr (HOT)
    tally_clear(P, lds_tally);
  const long long base = a.inst_start[inst];
  const long long count = a.inst_count[inst];
  const long long first = (long long)tile * TS - base + threadIdx.x;  // local index, j = 0

  R kx[PPT], ky[PPT], kz[PPT], E[PPT], tau[PPT], rem[PPT];
  R pos[PPT][3];  // positions (POS only; folded away otherwise)
  int v[PPT], sub[PPT];
  uint32_t ev[PPT];
  bool valid[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const long long local = first + (long long)j * kTile;
    valid[j] = local < count;
    kx[j] = ky[j] = kz[j] = E[j] = tau[j] = rem[j] = R(0);
    pos[j][0] = pos[j][1] = pos[j][2] = R(0);
    v[j] = sub[j] = 0;
    ev[j] = 0;
    if (valid[j]) {
      const long long gi = base + local;
      kx[j] = a.kx[gi];
      ky[j] = a.ky[gi];
      kz[j] = a.kz[gi];
      E[j] = a.E[gi];
      tau[j] = a.tau[gi];
      const uint32_t vs = a.vs[gi];
      v[j] = (int)(vs >> 16);
      sub[j] = (int)(vs & 0xffffu);
      if constexpr (POS) {
        pos[j][0] = a.px[gi];
        pos[j][1] = a.py[gi];
        pos[j][2] = a.pz[gi];
      }
    }
  }
  StepRng<R, 0> rng;
  rng.inst = (uint32_t)inst;
  rng.k0 = a.key0;
  rng.k1 = a.key1;
  const R dt = P->dt;

  for (int s = 0; s < a.n_steps; ++s) {
    rng.step = a.step0 + (uint32_t)s;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      ev[j] = 0;
      rem[j] = R(0);
      if (valid[j]) {
        drift<R, POS>(P, (dt < tau[j]) ? dt : tau[j], kx[j], ky[j], kz[j], E[j], v[j], sub[j],
                      pos[j]);
        rem[j] = dt - tau[j];  // std::min(tau, dt) above, then tRemaining
      }
    }
    int nreal = 0;
    for (;;) {
      // next particle of this lane with time left in the step (lowest j)
      int sel = -1;
#pragma unroll
      for (int j = PPT - 1; j >= 0; --j)
        if (valid[j] && rem[j] > R(0) && ev[j] < kMaxEventsPerStep)
          sel = j;
      if (sel < 0)
        break;
      R x = kx[0], y = ky[0], z = kz[0], en = E[0], ta = tau[0], rm = rem[0];
      R wp[3] = {pos[0][0], pos[0][1], pos[0][2]};
      int vv = v[0], ss = sub[0];
      uint32_t ee = ev[0];
#pragma unroll
      for (int j = 1; j < PPT; ++j)
        if (sel == j) {
          x = kx[j];
          y = ky[j];
          z = kz[j];
          en = E[j];
          ta = tau[j];
          rm = rem[j];
          vv = v[j];
          ss = sub[j];
          ee = ev[j];
          if constexpr (POS) {
            wp[0] = pos[j][0];
            wp[1] = pos[j][1];
            wp[2] = pos[j][2];
          }
        }
      rng.pid = (uint32_t)(first + (long long)sel * kTile);
      event_once<R, 0, HOT, HOT, POS>(P, a.cum, x, y, z, en, ta, rm, vv, ss, ee, rng, nreal,
                                      lds_tally, wp);
#pragma unroll
      for (int j = 0; j < PPT; ++j)
        if (sel == j) {
          kx[j] = x;
          ky[j] = y;
          kz[j] = z;
          E[j] = en;
          tau[j] = ta;
          rem[j] = rm;
          v[j] = vv;
          sub[j] = ss;
          ev[j] = ee;
          if constexpr (POS) {
            pos[j][0] = wp[0];
            pos[j][1] = wp[1];
            pos[j][2] = wp[2];
          }
        }
    }
    int nev = 0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      nev += (int)ev[j];
      if (valid[j])
        tau[j] -= dt;
    }
    tile_observables<R, PPT>(a, P, tile, s, valid, v, sub, kx, ky, kz, E, nev, nreal);
  }
  if constexpr (HOT)
    tally_flush(P, lds_tally, a.tally, a.tally_bins);

#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    if (valid[j]) {
      const long long gi = base + first + (long long)j * kTile;
      a.kx[gi] = kx[j];
      a.ky[gi] = ky[j];
      a.kz[gi] = kz[j];
      a.E[gi] = E[j];
      a.tau[gi] = tau[j];
      a.vs[gi] = ((uint32_t)v[j] << 16) | (uint32_t)sub[j];
      if constexpr (POS) {
        a.px[gi] = pos[j][0];
        a.py[gi] = pos[j][1];
        a.pz[gi] = pos[j][2];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// replay step kernel (FP64): one particle per lane, the reference engine's
// recorded draws consumed in the reference's order, optional per-step trace
// ---------------------------------------------------------------------------
template <typename R, bool HOT, bool POS>
__global__ __launch_bounds__(kTile) void k_bulk_replay(StepArgs<R> a) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds_tally[];
  const int tile = blockIdx.x;
  const int inst = tile_instance(a.unit_inst, tile, kTile);
  const InstDev<R> *P = a.inst + inst;
  if constexpr (HOT)
    tally_clear(P, lds_tally);
  const long long base = a.inst_start[inst];
  const long long local = (long long)tile * kTile + threadIdx.x - base;
  bool valid[1] = {local < a.inst_count[inst]};
  const long long gi = base + local;

  R kx[1] = {0}, ky[1] = {0}, kz[1] = {0}, E[1] = {0}, tau = 0;
  R pos[3] = {0, 0, 0};
  int v[1] = {0}, sub[1] = {0};
  if (valid[0]) {
    kx[0] = a.kx[gi];
    ky[0] = a.ky[gi];
    kz[0] = a.kz[gi];
    E[0] = a.E[gi];
    tau = a.tau[gi];
    const uint32_t vs = a.vs[gi];
    v[0] = (int)(vs >> 16);
    sub[0] = (int)(vs & 0xffffu);
    if constexpr (POS) {
      pos[0] = a.px[gi];
      pos[1] = a.py[gi];
      pos[2] = a.pz[gi];
    }
  }
  StepRng<R, 1> rng;
  rng.rs.p = a.draws + (valid[0] ? local : 0) * a.draw_stride;
  rng.rs.len = a.draw_stride;
  rng.rs.pos = valid[0] ? (long long)a.cursor[local] : 0;
  rng.rs.over = false;
  double grain = (valid[0] && a.grain_tau) ? a.grain_tau[local] : 1.0;

  for (int s = 0; s < a.n_steps; ++s) {
    int nev = 0, nreal = 0;
    if (valid[0]) {
      move_one<R, 1, HOT, HOT, POS>(P, a.cum, kx[0], ky[0], kz[0], E[0], tau, v[0], sub[0], rng,
                                    nev, nreal, lds_tally, pos);
      // grain bookkeeping (basicBulkParticleHandler.hpp:213-220), no grain
      // mechanism: a new grainTau of -ln U[1e-6,1) * 1 s
      grain -= a.dt_s;
      if (grain <= 0)
        grain = -::log(rng.rs.uniform(1e-6, 1.)) * 1.0;
      if (a.tr_valley && inst == 0) {
        const long long o = (long long)s * a.tr_n + local;
        a.tr_valley[o] = v[0];
        a.tr_sub[o] = sub[0];
        a.tr_events[o] = nev;
        a.tr_k[3 * o] = (double)kx[0];
        a.tr_k[3 * o + 1] = (double)ky[0];
        a.tr_k[3 * o + 2] = (double)kz[0];
        a.tr_E[o] = (double)E[0];
        if (a.tr_pos) {
          a.tr_pos[3 * o] = (double)pos[0];
          a.tr_pos[3 * o + 1] = (double)pos[1];
          a.tr_pos[3 * o + 2] = (double)pos[2];
        }
      }
    }
    tile_observables<R, 1>(a, P, tile, s, valid, v, sub, kx, ky, kz, E, nev, nreal);
  }
  if constexpr (HOT)
    tally_flush(P, lds_tally, a.tally, a.tally_bins);

  if (valid[0]) {
    a.kx[gi] = kx[0];
    a.ky[gi] = ky[0];
    a.kz[gi] = kz[0];
    a.E[gi] = E[0];
    a.tau[gi] = tau;
    a.vs[gi] = ((uint32_t)v[0] << 16) | (uint32_t)sub[0];
    if constexpr (POS) {
      a.px[gi] = pos[0];
      a.py[gi] = pos[1];
      a.pz[gi] = pos[2];
    }
    a.cursor[local] = (uint64_t)rng.rs.pos;
    if (a.grain_tau)
      a.grain_tau[local] = grain;
    if (rng.rs.over)
      atomicAdd(a.overflow, 1);
  }
}

// the ensemble sums of the current state, without a step: one particle per
// lane, the partials of "step 0" (the getAvg* calls right after
// generateInitialParticles, bulkSimulation.cpp:143-146)
template <typename R>
__global__ __launch_bounds__(kTile) void k_observe(StepArgs<R> a) {
  const int tile = blockIdx.x;
  const int inst = tile_instance(a.unit_inst, tile, kTile);
  const InstDev<R> *P = a.inst + inst;
  const long long base = a.inst_start[inst];
  const long long local = (long long)tile * kTile + threadIdx.x - base;
  bool valid[1] = {local < a.inst_count[inst]};
  const long long gi = base + local;
  R kx[1] = {0}, ky[1] = {0}, kz[1] = {0}, E[1] = {0};
  int v[1] = {0}, sub[1] = {0};
  if (valid[0]) {
    kx[0] = a.kx[gi];
    ky[0] = a.ky[gi];
    kz[0] = a.kz[gi];
    E[0] = a.E[gi];
    const uint32_t vs = a.vs[gi];
    v[0] = (int)(vs >> 16);
    sub[0] = (int)(vs & 0xffffu);
  }
  tile_observables<R, 1>(a, P, tile, 0, valid, v, sub, kx, ky, kz, E, 0, 0);
}

// ---------------------------------------------------------------------------
// per-(step, observable, instance) compensated reduction of the tile partials
// partial layout [step][obs][tile].  Two passes, both deterministic (fixed
// order for a given tile count): chunks of kChunk tiles per block with
// Neumaier summation per thread and a TwoSum tree across the block, then
// the chunk (sum, compensation) pairs of each output value.
// ---------------------------------------------------------------------------
constexpr int kChunk = 2048;

__device__ __forceinline__ void neumaier_add(double &sum, double &comp, double x) {
  const double y = sum + x;
  if (fabs(sum) >= fabs(x))
    comp += (sum - y) + x;
  else
    comp += (x - y) + sum;
  sum = y;
}

// TwoSum tree over 256 (sum, comp) pairs in LDS; result in ss[0] + sc[0]
__device__ __forceinline__ void block_twosum(double *ss, double *sc) {
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double x = ss[threadIdx.x], y0 = ss[threadIdx.x + w];
      const double y = x + y0;
      const double bb = y - x;
      const double err = (x - (y - bb)) + (y0 - bb);
      ss[threadIdx.x] = y;
      sc[threadIdx.x] += sc[threadIdx.x + w] + err;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_reduce_chunks(const double *__restrict__ partial,
                                                       int n_tiles, int tile_size,
                                                       const long long *__restrict__ inst_start,
                                                       const long long *__restrict__ inst_count,
                                                       int n_chunks, double *__restrict__ chunk_out) {
  __shared__ double ss[256], sc[256];
  const int chunk = blockIdx.x, so = blockIdx.y, inst = blockIdx.z;
  // the wave partials of the instance's tiles (tiles start on instance
  // boundaries; every wave of a tile with particles wrote one)
  const long long t0 = inst_start[inst] / tile_size * kSubs;
  const long long nt = (inst_count[inst] + tile_size - 1) / tile_size * kSubs;
  const long long lo = (long long)chunk * kChunk;
  const long long hi = (lo + kChunk < nt) ? lo + kChunk : nt;
  const double *src = partial + (long long)so * ((long long)n_tiles * kSubs) + t0;
  double sum = 0.0, comp = 0.0;
  for (long long t = lo + threadIdx.x; t < hi; t += 256)
    neumaier_add(sum, comp, src[t]);
  ss[threadIdx.x] = sum;
  sc[threadIdx.x] = comp;
  block_twosum(ss, sc);
  if (threadIdx.x == 0) {
    double *o = chunk_out + (((long long)inst * gridDim.y + so) * n_chunks + chunk) * 2;
    o[0] = ss[0];
    o[1] = sc[0];
  }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double *__restrict__ chunk_out,
                                                      int n_chunks, int n_obs,
                                                      double *__restrict__ out, int out_step0,
                                                      int out_total) {
  __shared__ double ss[256], sc[256];
  const int so = blockIdx.x, inst = blockIdx.y;
  const int s = so / n_obs, o = so % n_obs;
  const double *src = chunk_out + ((long long)inst * gridDim.x + so) * n_chunks * 2;
  double sum = 0.0, comp = 0.0;
  for (int c = threadIdx.x; c < n_chunks; c += 256) {
    neumaier_add(sum, comp, src[2 * c]);
    comp += src[2 * c + 1];
  }
  ss[threadIdx.x] = sum;
  sc[threadIdx.x] = comp;
  block_twosum(ss, sc);
  if (threadIdx.x == 0)
    out[((long long)inst * out_total + out_step0 + s) * n_obs + o] = ss[0] + sc[0];
}

// ---------------------------------------------------------------------------
// End of a hot-phonon step, one block per instance with baths:
//   every bath of the instance is updated (emcPhononBath::update,
//   emcPhononBath.hpp:264-358) and its mean occupation taken (getMeanNq,
//   :461-470); then, on rebuild steps, the instance's scatter tables are
//   rebuilt from the baths (reinitScatterTables: fillScatterTables +
//   renormalizeTables, emcScatterHandler.hpp:100-108, 220-272).
// The per-bin update runs on all lanes; the DOS-weighted sums that the
// reference takes sequentially over the bins are taken in that order by
// lane 0 from LDS, so the FP64 results match the reference operation for
// operation.  Hot-phonon rates are scatterConst * N / kI * ln (absorption)
// and scatterConst * (N + 1) / kI * ln (emission) with the per-level kI and
// ln of emcHotPhononFroehlichMechanism.hpp:79-91 / :159-173 precomputed in
// FP64 on the host (kI = 0 marks a zero rate); other mechanisms keep their
// FP64 rates.  Cumulative sums run over the tables in registration order and
// are normalised by the largest total rate; the FP64 result is stored in the
// precision of the kinematics and in an FP64 copy.
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;
constexpr double kQ = 1.60219e-19, kKB = 1.38066e-23;  // include/emcConstants.hpp
constexpr double kEps0 = 8.85419e-12;
constexpr int kBathThreads = 1024;

__device__ double bath_planck(double energy_ev, double temp_k) {
#pragma clang fp contract(off)
  const double x = kQ * energy_ev / (kKB * temp_k);
  return 1.0 / (::exp(x) - 1.0);
}
__device__ double bath_planck_temp(double energy_ev, double n) {
#pragma clang fp contract(off)
  if (n <= 0.0)
    return 0.0;
  return kQ * energy_ev / (kKB * ::log(1.0 + 1.0 / n));
}

// one lane per bin in the bath update; the per-bin LDS arrays double as
// per-lane reduction buffers
static_assert(kMaxBins == kBathThreads, "bath block: one lane per bin");

struct BathShared {
  double old_n[kMaxBins];
  double new_n[kMaxBins];
  double w[kMaxBins];
  double vmax[kBathThreads];
  // terms of the energy record: net emissions, LO population change, decay
  double e_em[kMaxBins], e_lo[kMaxBins], e_dec[kMaxBins];
  double target;
  double qs2;  // the instance's q_s^2, broadcast to the block
};

template <typename R>
__device__ void bath_update_one(const BathDev &B, BathVar &V, double *__restrict__ nq,
                                const double *__restrict__ tau_prof,
                                unsigned long long *__restrict__ tally, long long tally_bins,
                                double dt, double *__restrict__ hist,
                                double *__restrict__ energy, BathShared &sh,
                                double qs2, double *__restrict__ win64, R *__restrict__ win) {
#pragma clang fp contract(off)
  const int nb = B.n_bins;
  if (threadIdx.x == 0) {
    // decay target: N0 unless a daughter reservoir is hot (:269-286)
    double target = B.n0;
    if (B.has_ac) {
      double n_klemens = B.n0;
      if (V.n_ac > B.n_ac_eq)
        n_klemens = bath_planck(B.lo_e, bath_planck_temp(B.ac_e, V.n_ac));
      if (B.has_ridley) {
        double n_ridley = B.n0;
        if (V.n_to > B.n_to_eq)
          n_ridley = bath_planck(B.lo_e, bath_planck_temp(B.to_e, V.n_to));
        target = (1.0 - B.w_r) * n_klemens + B.w_r * n_ridley;
      } else {
        target = n_klemens;
      }
    }
    sh.target = target;
  }
  // the event counts of this step: the sum of the tally copies, which are
  // cleared.  G = blockDim / nb lane groups share the copies of a bin
  // (copies g, g + G, ...), their integer partial sums meet in LDS (exact as
  // doubles: far below 2^53)
  const int G = (nb > 0 && (int)blockDim.x >= nb) ? (int)blockDim.x / nb : 1;
  if ((int)threadIdx.x < G * nb) {
    const int i = (int)threadIdx.x % nb, g = (int)threadIdx.x / nb;
    const long long gi = B.off + i;
    unsigned long long c_em = 0ull, c_abs = 0ull;
#pragma unroll 16
    for (int r = g; r < kTallyCopies; r += G) {
      unsigned long long *t = tally + 2 * ((long long)r * tally_bins + gi);
      c_em += t[0];
      c_abs += t[1];
      t[0] = 0ull;
      t[1] = 0ull;
    }
    sh.vmax[threadIdx.x] = (double)c_em;
    sh.old_n[threadIdx.x] = (double)c_abs;
  }
  __syncthreads();
  double cnt_em[kMaxBins / kBathThreads + 1], cnt_abs[kMaxBins / kBathThreads + 1];
  {
    int u = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x, ++u) {
      double e = 0.0, b = 0.0;
      for (int g = 0; g < G; ++g) {
        e = e + sh.vmax[g * nb + i];
        b = b + sh.old_n[g * nb + i];
      }
      cnt_em[u] = e;
      cnt_abs[u] = b;
    }
  }
  __syncthreads();
  const double target = sh.target;
  int u = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x, ++u) {
    const long long gi = B.off + i;
    const double q = ((double)i + 0.5) * B.dq;
    const double dph = q * q * B.dq * B.vsim / (2.0 * kPi * kPi);
    const double n_em = cnt_em[u], n_abs = cnt_abs[u];
    const double g = (dph > 0.0) ? (n_em - n_abs) / (dph * dt) : 0.0;
    const double tau_i = B.has_profile ? tau_prof[gi] : B.tau_lo;
    const double n_old = nq[gi];
    const double w = q * q;
    double n = n_old;
    n += g * dt - (dt / tau_i) * (n - target);
    if (n < 0.0)
      n = 0.0;
    nq[gi] = n;
    if (energy) {
      sh.e_em[i] = n_em - n_abs;
      sh.e_lo[i] = dph * (n - n_old);
      sh.e_dec[i] = dph * (dt / tau_i) * (n_old - target);
    }
    // the terms of the bin-order sums, formed on all lanes: the reservoir
    // source from the start-of-step occupations (:291-355), getMeanNq
    // (:461-470) and, when read, the window sums (:109-124)
    sh.old_n[i] = B.has_profile ? w * (n_old - target) / tau_i : w * (n_old - target);
    sh.new_n[i] = w * n;
    if (win64) {
      const double denom = q * q + qs2;
      const double ww = (denom > 0.0) ? q / denom : 0.0;
      sh.w[i] = ww;
      sh.vmax[i] = ww * n;
    }
  }
  __syncthreads();
  // the sums themselves in bin order (the reference's), one lane each of
  // four waves side by side
  const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  if (lane == 0 && wave == 0 && B.has_ac) {
    {
      // sum of the weights q^2: a constant of the bath, summed on the host
      // in the same order
      const double sum_w = B.sum_w;
      double sum_t = 0.0;
#pragma unroll 16
      for (int i = 0; i < nb; ++i)
        sum_t += sh.old_n[i];
      const double f_klemens = B.has_ridley ? (1.0 - B.w_r) : 1.0;
      const double src = (sum_w > 0.0) ? sum_t / sum_w : 0.0;
      double n_ac = V.n_ac, n_to = V.n_to;
      if (!B.has_profile) {
        n_ac += dt * f_klemens * src / B.tau_lo - dt * (n_ac - B.n_ac_eq) / B.tau_ac;
        if (n_ac < B.n_ac_eq)
          n_ac = B.n_ac_eq;
        if (B.has_ridley) {
          n_to += dt * B.w_r * src / B.tau_lo - dt * (n_to - B.n_to_eq) / B.tau_to;
          if (n_to < B.n_to_eq)
            n_to = B.n_to_eq;
        }
      } else {
        n_ac += dt * f_klemens * src - dt * (n_ac - B.n_ac_eq) / B.tau_ac;
        if (n_ac < B.n_ac_eq)
          n_ac = B.n_ac_eq;
        if (B.has_ridley) {
          n_to += dt * B.w_r * src - dt * (n_to - B.n_to_eq) / B.tau_to;
          if (n_to < B.n_to_eq)
            n_to = B.n_to_eq;
        }
      }
      V.n_ac = n_ac;
      V.n_to = n_to;
    }
  } else if (lane == 0 && wave == 1) {
    // getMeanNq over the updated occupations
    double sum_wn = 0.0;
#pragma unroll 16
    for (int i = 0; i < nb; ++i)
      sum_wn += sh.new_n[i];
    const double mean = (B.sum_w > 0.0) ? sum_wn / B.sum_w : B.n0;
    V.mean = mean;
    if (hist)
      *hist = mean;
  } else if (lane == 0 && wave == 2 && win64) {
    double *cw = win64 + B.woff;
    R *rw = win + B.woff;
    double acc = 0.0;
    cw[0] = 0.0;
    rw[0] = R(0);
#pragma unroll 16
    for (int i = 0; i < nb; ++i) {
      acc = acc + sh.w[i];
      cw[i + 1] = acc;
      rw[i + 1] = (R)acc;
    }
  } else if (lane == 0 && wave == 3 && win64) {
    double *cwn = win64 + B.woff + nb + 1;
    R *rwn = win + B.woff + nb + 1;
    double acc = 0.0;
    cwn[0] = 0.0;
    rwn[0] = R(0);
#pragma unroll 16
    for (int i = 0; i < nb; ++i) {
      acc = acc + sh.vmax[i];
      cwn[i + 1] = acc;
      rwn[i + 1] = (R)acc;
    }
  } else if (lane == 0 && wave >= 4 && wave <= 6 && energy) {
    // the energy record in bin order, x hbar w_LO [eV]
    const double *t = (wave == 4) ? sh.e_em : (wave == 5) ? sh.e_lo : sh.e_dec;
    double acc = 0.0;
#pragma unroll 16
    for (int i = 0; i < nb; ++i)
      acc += t[i];
    energy[wave - 4] = B.lo_e * acc;
  }
  __syncthreads();
}

// getNqInWindow (emcPhononBath.hpp:378-394) over the FP64 window sums
__device__ __forceinline__ double window_nq(const BathDev &B, const double *__restrict__ win64,
                                           double mean, double qmin, double qmax) {
#pragma clang fp contract(off)
  if (B.n_bins < 2 || qmax <= qmin)
    return mean;
  long lo = (long)::floor(qmin / B.dq);
  long hi = (long)::ceil(qmax / B.dq);
  if (lo < 0)
    lo = 0;
  if (hi > (long)B.n_bins)
    hi = (long)B.n_bins;
  if (hi - lo < 1)
    return mean;
  const double *cw = win64 + B.woff, *cwn = cw + B.n_bins + 1;
  const double wsum = cw[hi] - cw[lo];
  if (wsum <= 0.0)
    return mean;
  return (cwn[hi] - cwn[lo]) / wsum;
}

// screenedFroehlichLogFactor (emcScreenedFroehlichInteraction.hpp:52-62)
__device__ __forceinline__ double screened_log(double ki, double kf, double qs2) {
#pragma clang fp contract(off)
  const double p2 = (ki + kf) * (ki + kf), m2 = (ki - kf) * (ki - kf);
  if (qs2 <= 0.0)
    return (m2 <= 0.0) ? 0.0 : 0.5 * ::log(p2 / m2);
  return 0.5 * ::log((p2 + qs2) / (m2 + qs2));
}

__device__ __forceinline__ double rebuild_rate(const RebuildJob &J, int j, int l,
                                               const double *__restrict__ arena,
                                               const BathVar *__restrict__ var,
                                               const BathDev *__restrict__ baths,
                                               const double *__restrict__ win64, double qs2) {
#pragma clang fp contract(off)
  const long long o = J.rate_off[j];
  if (J.kind[j] == kRateConst)
    return arena[o + l];
  const double ki = arena[o + l];
  if (!(ki > 0.0))
    return 0.0;
  if (J.kind[j] == kRateScrAb || J.kind[j] == kRateScrEm) {
    // emcScreenedFroehlichInteraction.hpp:128-138 / :184-196 (constant N)
    // and :257-269 / :333-348 (bath mean or window mean)
    const double kf = arena[o + J.n_levels + l];
    double n = J.nconst[j];
    if (J.nsrc[j] == kNMean) {
      n = var[J.bath[j]].mean;
    } else if (J.nsrc[j] == kNWindow) {
      const double d = ki - kf;
      n = window_nq(baths[J.bath[j]], win64, var[J.bath[j]].mean, (d < 0.0) ? -d : d, ki + kf);
    }
    const double lf = screened_log(ki, kf, qs2);
    return (J.kind[j] == kRateScrAb) ? J.C[j] * n / ki * lf : J.C[j] * (n + 1.0) / ki * lf;
  }
  const double ln = arena[o + J.n_levels + l];
  const double n = var[J.bath[j]].mean;
  return (J.kind[j] == kRateHotAb) ? J.C[j] * n / ki * ln : J.C[j] * (n + 1.0) / ki * ln;
}

template <typename R>
__device__ void rebuild_one(const RebuildJob &J, const BathStepArgs<R> &a, double qs2,
                            BathShared &sh) {
#pragma clang fp contract(off)
  const double *__restrict__ arena = a.arena;
  const BathVar *__restrict__ var = a.var;
  const BathDev *__restrict__ baths = a.baths;
  const double *__restrict__ win64 = a.win64;
  R *__restrict__ cum = a.cum;
  double *__restrict__ cum64 = a.cum64;
  const int nt = J.n_tables, nl = J.n_levels;
  // pass 1: the cumulative rates of each level, kept unnormalised in the
  // FP64 copy by the lane that normalises them in pass 2
  double mx = -1.0;
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    double c = 0.0;
    for (int j = 0; j < nt; ++j) {
      const double r = rebuild_rate(J, j, l, arena, var, baths, win64, qs2);
      c = (j == 0) ? r : r + c;
      cum64[J.cum_off + (long long)l * nt + j] = c;
    }
    mx = (mx < c) ? c : mx;
  }
  sh.vmax[threadIdx.x] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && sh.vmax[threadIdx.x] < sh.vmax[threadIdx.x + w])
      sh.vmax[threadIdx.x] = sh.vmax[threadIdx.x + w];
    __syncthreads();
  }
  mx = sh.vmax[0];
  __syncthreads();
  if (!(mx > 0.0))
    return;  // no rate at any level: the kinematics keep the previous tables
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    for (int j = 0; j < nt; ++j) {
      const long long e = J.cum_off + (long long)l * nt + j;
      const double x = cum64[e] / mx;
      cum[e] = (R)x;
      cum64[e] = x;
    }
  }
  if (threadIdx.x == 0) {
    const double tau = 1. / mx;
    a.inst[J.inst].tau0[J.valley] = (R)(tau * 1e15);
    a.tau0_64[J.inst * kMaxV + J.valley] = tau;
  }
}

// updateScreening (examples/hotCarrierMHP/hotCarrierMHP.cpp:563-581) of one
// instance from the partials of the step just taken: N over all valleys,
// the valley-0 energy sum and count, the Debye q_s^2 at the temperature of
// the valley-0 mean energy (getMBTemp, emcHotCarrierOutput.hpp:89-92)
template <typename R>
__device__ void screening_update(const BathInstDev &I, const BathStepArgs<R> &a, BathShared &sh) {
#pragma clang fp contract(off)
  double vals[3] = {0.0, 0.0, 0.0};  // N, sum E valley 0, count valley 0
  // the instance's tiles: it starts on a tile boundary, tiles past its
  // count hold zeros
  const long long ts = a.tile_size, nsub = (long long)a.n_tiles * kSubs;
  const long long tile0 = a.inst_start[I.inst] / ts * kSubs;
  const long long tile1 = (a.inst_start[I.inst] + a.inst_count[I.inst] + ts - 1) / ts * kSubs;
  const long long tend = (tile1 < nsub) ? tile1 : nsub;
#pragma unroll 4
  for (long long t = tile0 + threadIdx.x; t < tend; t += blockDim.x) {
    for (int v = 0; v < I.n_valleys; ++v)
      vals[0] += a.partial[(long long)(2 + 3 * v) * nsub + t];
    vals[1] += a.partial[3LL * nsub + t];
    vals[2] += a.partial[2LL * nsub + t];
  }
  // the three sums in one tree (fixed order, deterministic)
  double *r0 = sh.vmax, *r1 = sh.old_n, *r2 = sh.new_n;
  r0[threadIdx.x] = vals[0];
  r1[threadIdx.x] = vals[1];
  r2[threadIdx.x] = vals[2];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r0[threadIdx.x] = r0[threadIdx.x] + r0[threadIdx.x + w];
      r1[threadIdx.x] = r1[threadIdx.x] + r1[threadIdx.x + w];
      r2[threadIdx.x] = r2[threadIdx.x] + r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  const double tot[3] = {r0[0], r1[0], r2[0]};
  __syncthreads();
  if (threadIdx.x == 0) {
    double qs2 = 0.0;
    if (tot[0] != 0.0) {
      const double ns = tot[0] / I.scr_vol;
      const double avg = (tot[2] != 0.0) ? tot[1] / tot[2] : 0.0;
      const double ts = 2.0 * avg * kQ / (3.0 * kKB);
      if (ts > 0.0)
        qs2 += ns * kQ * kQ / (I.scr_eps * kEps0 * kKB * ts);
    }
    a.scr[I.inst] = qs2;
    a.inst[I.inst].qs2 = (R)(qs2 * 1e-18);
    sh.qs2 = qs2;
  }
  __syncthreads();
}

// rebuildWindowSums (emcPhononBath.hpp:109-124) of one bath with the
// instance's q_s^2: weights on all lanes, the prefix sums by lane 0 in bin
// order (FP64, as the reference adds them) and a copy in the precision of
// the kinematics
template <typename R>
__device__ void window_sums_one(const BathDev &B, double qs2, const BathStepArgs<R> &a,
                                BathShared &sh) {
#pragma clang fp contract(off)
  const int nb = B.n_bins;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const double q = ((double)i + 0.5) * B.dq;
    const double denom = q * q + qs2;
    const double w = (denom > 0.0) ? q / denom : 0.0;
    sh.old_n[i] = w;
    sh.new_n[i] = w * a.nq[B.off + i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *cw = a.win64 + B.woff, *cwn = cw + nb + 1;
    R *rw = a.win + B.woff, *rwn = rw + nb + 1;
    double s = 0.0, sn = 0.0;
    cw[0] = 0.0;
    cwn[0] = 0.0;
    rw[0] = R(0);
    rwn[0] = R(0);
#pragma unroll 16
    for (int i = 0; i < nb; ++i) {
      s = s + sh.old_n[i];
      sn = sn + sh.new_n[i];
      cw[i + 1] = s;
      cwn[i + 1] = sn;
      rw[i + 1] = (R)s;
      rwn[i + 1] = (R)sn;
    }
  }
  __syncthreads();
}

// the instance's q_s^2 into LDS: from the step's partials (plasma
// screening) or as held
template <typename R>
__device__ void group_screening(const BathInstDev &I, const BathStepArgs<R> &a, bool plasma,
                                BathShared &sh) {
  if (plasma && I.plasma) {
    screening_update(I, a, sh);
    return;
  }
  if (threadIdx.x == 0)
    sh.qs2 = a.scr ? a.scr[I.inst] : 0.0;
  __syncthreads();
}

template <typename R>
__device__ void rebuild_group(const BathInstDev &I, const BathStepArgs<R> &a, BathShared &sh) {
  const double qs2 = sh.qs2;
  for (int j = I.job0; j < I.job0 + I.n_jobs; ++j)
    rebuild_one<R>(a.jobs[j], a, qs2, sh);
}

// One block per instance with baths or plasma screening: the baths' update,
// the screening from the step's partials, the window sums, the tables (the
// order of hotCarrierMHP.cpp:693-703).  Window sums and screening follow
// the values the reference's setScreeningQ2 / update leave behind.
// The screening depends on the carriers only, so it is taken first; the
// window sums then see the updated occupations and the new q_s^2, as after
// the reference's update + setScreeningQ2.
template <typename R>
__global__ __launch_bounds__(kBathThreads) void k_bath_step(BathStepArgs<R> a) {
  __shared__ BathShared sh;
  const BathInstDev &I = a.groups[blockIdx.x];
  group_screening<R>(I, a, a.partial != nullptr, sh);
  const double qs2 = sh.qs2;
  double *win64 = (a.win64 && I.need_win) ? a.win64 : nullptr;
  for (int b = I.bath0; b < I.bath0 + I.n_baths; ++b)
    bath_update_one<R>(a.baths[b], a.var[b], a.nq, a.tau_prof, a.tally, a.tally_bins, a.dt,
                       a.hist ? a.hist + (long long)b * a.hist_steps + a.hist_step : nullptr,
                       a.energy ? a.energy + 3 * ((long long)b * a.hist_steps + a.hist_step)
                                : nullptr,
                       sh, qs2, win64, a.win);
  // the tables follow in k_rebuild_rates / k_rebuild_norm
}

// The per-step table rebuild over (level chunk, job) blocks, one level per
// lane: the cumulative FP64 rates of the level (kept unnormalised in the
// FP64 copy) and the block's maximum ...
template <typename R>
__global__ __launch_bounds__(kRebuildChunk) void k_rebuild_rates(BathStepArgs<R> a) {
#pragma clang fp contract(off)
  __shared__ double red[kRebuildChunk];
  const RebuildJob &J = a.jobs[blockIdx.y];
  if (!(a.rebuild || J.every_step))
    return;
  const int nl = J.n_levels, nt = J.n_tables;
  const int l0 = (int)blockIdx.x * kRebuildChunk;
  if (l0 >= nl)
    return;  // the whole block leaves: no barrier below is reached
  const double qs2 = a.scr ? a.scr[J.inst] : 0.0;
  const int l = l0 + (int)threadIdx.x;
  double c = -1.0;
  if (l < nl) {
    for (int j = 0; j < nt; ++j) {
      const double r = rebuild_rate(J, j, l, a.arena, a.var, a.baths, a.win64, qs2);
      c = (j == 0) ? r : r + c;
      a.cum64[J.cum_off + (long long)l * nt + j] = c;
    }
  }
  red[threadIdx.x] = c;
  __syncthreads();
  for (int w = kRebuildChunk / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && red[threadIdx.x] < red[threadIdx.x + w])
      red[threadIdx.x] = red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    a.rb_max[(long long)blockIdx.y * kRebuildChunks + blockIdx.x] = red[0];
}

// ... then the largest total rate of the job over its chunks, the
// normalised tables and tau0 (renormalizeTables, emcScatterHandler.hpp:
// 248-272); with no rate at any level the kinematics keep the previous tables
template <typename R>
__global__ __launch_bounds__(kRebuildChunk) void k_rebuild_norm(BathStepArgs<R> a) {
#pragma clang fp contract(off)
  const RebuildJob &J = a.jobs[blockIdx.y];
  if (!(a.rebuild || J.every_step))
    return;
  const int nl = J.n_levels, nt = J.n_tables;
  const int l0 = (int)blockIdx.x * kRebuildChunk;
  if (l0 >= nl)
    return;
  const int nc = (nl + kRebuildChunk - 1) / kRebuildChunk;
  double mx = -1.0;
  for (int k = 0; k < nc; ++k) {
    const double m = a.rb_max[(long long)blockIdx.y * kRebuildChunks + k];
    mx = (mx < m) ? m : mx;
  }
  if (!(mx > 0.0))
    return;
  const int l = l0 + (int)threadIdx.x;
  if (l < nl) {
    for (int j = 0; j < nt; ++j) {
      const long long e = J.cum_off + (long long)l * nt + j;
      const double x = a.cum64[e] / mx;
      a.cum[e] = (R)x;
      a.cum64[e] = x;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double tau = 1. / mx;
    a.inst[J.inst].tau0[J.valley] = (R)(tau * 1e15);
    a.tau0_64[J.inst * kMaxV + J.valley] = tau;
  }
}

// window sums and tables alone (after the model, the field, a bath state or
// the screening changed)
template <typename R>
__global__ __launch_bounds__(kBathThreads) void k_bath_refresh(BathStepArgs<R> a) {
  __shared__ BathShared sh;
  const BathInstDev &I = a.groups[blockIdx.x];
  group_screening<R>(I, a, false, sh);
  if (a.win64 && I.need_win)
    for (int b = I.bath0; b < I.bath0 + I.n_baths; ++b)
      window_sums_one(a.baths[b], sh.qs2, a, sh);
  rebuild_group<R>(I, a, sh);
}

// ---------------------------------------------------------------------------
// scatter rates from descriptors: one lane per (level, mechanism), the
// FP64 formulas of emcgpuPhysics.hpp (the reference's operation order; the
// device's pow / sqrt / log are within an ulp of the host's)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rates_desc(const RateDescJob *__restrict__ jobs,
                                                    int n_levels, double dE,
                                                    double *__restrict__ out) {
#pragma clang fp contract(off)
  const RateDescJob &J = jobs[blockIdx.y];
  const int l = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (l >= n_levels)
    return;
  const double E = (double)(unsigned long long)(l + 1) * dE;
  const bool ab = J.absorption != 0;
  double r = 0.0;
  switch (J.kind) {
  case EMCGPU_MECH_ACOUSTIC:
    r = phys::acousticRate(J.C, J.iv, E);
    break;
  case EMCGPU_MECH_ZERO_IV_AB:
  case EMCGPU_MECH_ZERO_IV_EM:
    r = phys::zeroOrderRate(J.C, J.fv, E, J.hw, J.delta, ab);
    break;
  case EMCGPU_MECH_FIRST_IV_AB:
  case EMCGPU_MECH_FIRST_IV_EM:
    r = phys::firstOrderRate(J.C, J.iv, J.fv, E, J.hw, J.delta, ab);
    break;
  case EMCGPU_MECH_SCREENED_FROEHLICH_AB:
  case EMCGPU_MECH_SCREENED_FROEHLICH_EM:
    r = phys::screenedFroehlichRate(J.C, J.nq, J.meff, J.iv, E, J.hw, J.qs2, ab);
    break;
  default:
    r = phys::froehlichRate(J.C, J.nq, J.meff, J.iv, E, J.hw, ab);
    break;
  }
  out[J.out_off + l] = r;
}

hipError_t launch_rates_desc(const RateDescJob *jobs, int n_jobs, int n_levels, double dE,
                             double *out, hipStream_t st) {
  if (n_jobs <= 0 || n_levels <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_rates_desc, dim3((n_levels + 255) / 256, n_jobs), dim3(256), 0, st, jobs,
                     n_levels, dE, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// thermal initialisation (emcElectron.hpp:75-90, emcParticleInitialization.hpp:36-51)
// ---------------------------------------------------------------------------
// init_mean > 0 selects the prescribed-energy initialiser: E = mean +
// sigma * sqrt(-2 ln u2) cos(2 pi u6) (u2 in [1e-6, 1)), |E| if negative,
// the mean if zero; sigma = 0 is the fixed-energy photoexcitation
// (emcParticleInitialization.hpp:57-71)
template <typename R>
__global__ __launch_bounds__(kTile) void k_init_thermal(StepArgs<R> a, double init_mean,
                                                        double init_sigma) {
  const long long tile = (long long)a.tile_base + blockIdx.x;
  const int inst = tile_instance(a.unit_inst, tile, kTile);
  const InstDev<R> *P = a.inst + inst;
  const long long base = a.inst_start[inst];
  const long long local = tile * kTile + threadIdx.x - base;
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
  R E;
  if (init_mean > 0.0) {
    E = (R)init_mean;
    if (init_sigma > 0.0) {
      const R u6 = stream_uniform<R>(6, st, pid, in, a.key0, a.key1);
      R sn, cs;
      Math<R>::sincos2pi(u6, &sn, &cs);
      E = (R)init_mean + (R)init_sigma * (Math<R>::sqrt(R(-2) * Math<R>::log(ue)) * cs);
      E = (E < R(0)) ? -E : E;
      E = (E == R(0)) ? (R)init_mean : E;
    }
  } else {
    E = R(-1.5) * P->vt * Math<R>::log(ue);
  }
  R kx, ky, kz;
  random_direction(normk_of(V, E), uphi, ucos, kx, ky, kz);
  a.kx[gi] = kx;
  a.ky[gi] = ky;
  a.kz[gi] = kz;
  a.E[gi] = E;
  a.tau[gi] = -Math<R>::log(ut) * P->tau0[v];
  a.vs[gi] = ((uint32_t)v << 16) | (uint32_t)sub;
  if (a.px) {
    // positions uniform in the periodic box: the density of the bulk
    // handler's initParticlePos per node (emcParticleInitialization.hpp:
    // 13-29) over nodes weighted by their cells
    a.px[gi] = stream_uniform<R>(8, st, pid, in, a.key0, a.key1) * P->box[0];
    a.py[gi] = stream_uniform<R>(9, st, pid, in, a.key0, a.key1) * P->box[1];
    a.pz[gi] = stream_uniform<R>(10, st, pid, in, a.key0, a.key1) * P->box[2];
  }
}

// ---------------------------------------------------------------------------
// launch wrappers (called by the C ABI layer)
// ---------------------------------------------------------------------------
// particles per tile of a step-kernel variant (1 = replay / one per lane)
int tile_size_of(int ppt) { return kTile * ppt; }

template <typename R, bool HOT>
hipError_t launch_steps_t(const StepArgs<R> &a, int mode, int ppt, hipStream_t st) {
  const dim3 g(a.n_tiles), b(kTile);
  // tallies, then (8-byte aligned) the staged window sums
  const size_t lds = HOT ? sizeof(uint32_t) * 2 * (size_t)a.lds_bins + 8 +
                               sizeof(R) * (size_t)a.lds_win
                         : 0;
  if (a.px) {
    // positions tracked: one particle per lane
    if (mode == 1)
      hipLaunchKernelGGL((k_bulk_replay<R, HOT, true>), g, b, lds, st, a);
    else if (ppt == 1)
      hipLaunchKernelGGL((k_bulk_multi<R, 1, HOT, true>), g, b, lds, st, a);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
  if (mode == 1) {
    hipLaunchKernelGGL((k_bulk_replay<R, HOT, false>), g, b, lds, st, a);
  } else {
    switch (ppt) {
    case 1:
      hipLaunchKernelGGL((k_bulk_multi<R, 1, HOT, false>), g, b, lds, st, a);
      break;
    case 2:
      hipLaunchKernelGGL((k_bulk_multi<R, 2, HOT, false>), g, b, lds, st, a);
      break;
    case 4:
      hipLaunchKernelGGL((k_bulk_multi<R, 4, HOT, false>), g, b, lds, st, a);
      break;
    case 8:
      hipLaunchKernelGGL((k_bulk_multi<R, 8, HOT, false>), g, b, lds, st, a);
      break;
    default:
      return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_steps(const StepArgs<R> &a, int mode, int ppt, hipStream_t st) {
  if (a.n_tiles <= 0)
    return hipSuccess;
  // the extended kernels: bath tallies and the screened final states
  if ((a.tally && a.lds_bins > 0) || a.ext_kinds)
    return launch_steps_t<R, true>(a, mode, ppt, st);
  return launch_steps_t<R, false>(a, mode, ppt, st);
}

template <typename R>
hipError_t launch_observe(const StepArgs<R> &a, hipStream_t st) {
  if (a.n_tiles <= 0)
    return hipSuccess;
  hipLaunchKernelGGL((k_observe<R>), dim3(a.n_tiles), dim3(kTile), 0, st, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_init(const StepArgs<R> &a, double init_mean, double init_sigma,
                       hipStream_t st) {
  if (a.n_tiles <= 0)
    return hipSuccess;
  hipLaunchKernelGGL((k_init_thermal<R>), dim3(a.n_tiles), dim3(kTile), 0, st, a, init_mean,
                     init_sigma);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_bath_step(const BathStepArgs<R> &a, int n_groups, hipStream_t st) {
  if (n_groups <= 0)
    return hipSuccess;
  hipLaunchKernelGGL((k_bath_step<R>), dim3(n_groups), dim3(kBathThreads), 0, st, a);
  if (a.n_jobs > 0) {
    const dim3 g(kRebuildChunks, a.n_jobs);
    hipLaunchKernelGGL((k_rebuild_rates<R>), g, dim3(kRebuildChunk), 0, st, a);
    hipLaunchKernelGGL((k_rebuild_norm<R>), g, dim3(kRebuildChunk), 0, st, a);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_rebuild(const BathStepArgs<R> &a, int n_jobs, hipStream_t st) {
  if (n_jobs <= 0)
    return hipSuccess;
  hipLaunchKernelGGL((k_bath_refresh<R>), dim3(n_jobs), dim3(kBathThreads), 0, st, a);
  return hipGetLastError();
}

// number of chunk pairs per output value for the largest instance
int reduce_chunks(long long max_count, int tile_size) {
  const long long nt = (max_count + tile_size - 1) / tile_size * kSubs;
  return (int)((nt + kChunk - 1) / kChunk > 0 ? (nt + kChunk - 1) / kChunk : 1);
}

hipError_t launch_reduce(const double *partial, int n_steps, int n_tiles, int n_obs, int tile_size,
                         const long long *inst_start, const long long *inst_count, int n_inst,
                         int n_chunks, double *chunk_buf, double *out, int out_step0,
                         int out_total, hipStream_t st) {
  if (n_steps <= 0 || n_inst <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_reduce_chunks, dim3(n_chunks, n_steps * n_obs, n_inst), dim3(256), 0, st,
                     partial, n_tiles, tile_size, inst_start, inst_count, n_chunks, chunk_buf);
  hipLaunchKernelGGL(k_reduce_final, dim3(n_steps * n_obs, n_inst), dim3(256), 0, st, chunk_buf,
                     n_chunks, n_obs, out, out_step0, out_total);
  return hipGetLastError();
}

template hipError_t launch_steps<float>(const StepArgs<float> &, int, int, hipStream_t);
template hipError_t launch_steps<double>(const StepArgs<double> &, int, int, hipStream_t);
template hipError_t launch_observe<float>(const StepArgs<float> &, hipStream_t);
template hipError_t launch_observe<double>(const StepArgs<double> &, hipStream_t);
template hipError_t launch_init<float>(const StepArgs<float> &, double, double, hipStream_t);
template hipError_t launch_init<double>(const StepArgs<double> &, double, double, hipStream_t);
template hipError_t launch_bath_step<float>(const BathStepArgs<float> &, int, hipStream_t);
template hipError_t launch_bath_step<double>(const BathStepArgs<double> &, int, hipStream_t);
template hipError_t launch_rebuild<float>(const BathStepArgs<float> &, int, hipStream_t);
template hipError_t launch_rebuild<double>(const BathStepArgs<double> &, int, hipStream_t);

}  // namespace emcgpu
