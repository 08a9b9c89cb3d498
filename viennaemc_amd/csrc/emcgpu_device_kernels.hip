// emcgpu_device_kernels.hip -- gfx950 kernels of the 2-D device simulation.
//
// One step of emcSimulation::performEMCStep (emcSimulation.hpp:164-193) is
// five launches on the context's stream:
//   k_dev_field         one workgroup: Poisson (Newton, FP64 residuals,
//                       FP32 red-black multigrid in LDS) and the E-field
//   k_dev_step          persistent blocks: drift/scatter of every particle
//                       (emcBasicParticleHandler.hpp:93-139), charge
//                       assignment into a block-private LDS grid, ranks of
//                       the particles at ohmic nodes in array order
//   k_dev_contact_scan  one block per ohmic node: array-order offsets of the
//                       segments, particles kept / deleted / to inject
//                       (handleOhmicContacts, emcBasicParticleHandler.hpp:
//                       158-192)
//   k_dev_resolve       deletions (holes) and injections (appended), with
//                       their charge corrections
//   k_dev_reduce        FP64 sum of the block grids -> carriers per node,
//                       concentration, commit of the ensemble counters
// Array order is the reference's: deleted particles leave holes (kHole),
// injected ones are appended, and a stable compaction between step chunks
// removes the holes.  Positions are in nm on the device.
#include <hip/hip_runtime.h>

#include <cstring>

#include "emcgpu_particle.h"

namespace emcgpu {

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
// posToCoord (emcDevice.hpp:290-297): round(pos / h) per dimension
template <typename R>
__device__ __forceinline__ int pos_coord(R p, R h) {
  return (int)Math<R>::round(Math<R>::div(p, h));
}

template <typename R>
__device__ __forceinline__ void acc_at(const DevGeom<R> &g, const R *acc, int ix, int iy, R &ax,
                                       R &ay) {
  if ((unsigned)ix < (unsigned)g.nx && (unsigned)iy < (unsigned)g.ny) {
    const R *q = acc + 2 * (iy * g.nx + ix);
    ax = q[0];
    ay = q[1];
  } else {
    ax = R(0);
    ay = R(0);
  }
}

// interpolateForce of the schemes on the force grid acc (g.acc, or a copy
// of it in LDS; nodes outside the grid count zero): NGP emcNGPScheme.hpp:
// 55-68, CIC emcCICScheme.hpp:122-141 with the (1 - wY)(1 - wY) weight of
// its last node, NEC emcNECScheme.hpp:99-113
template <typename R>
__device__ __forceinline__ void interp(const DevGeom<R> &g, const R *acc, R x, R y, R &fx,
                                       R &fy) {
#pragma clang fp contract(off)
  if (g.pm == EMCGPU_PM_NGP) {
    acc_at(g, acc, pos_coord(x, g.hx), pos_coord(y, g.hy), fx, fy);
    return;
  }
  const R sx = Math<R>::div(x, g.hx), sy = Math<R>::div(y, g.hy);
  const R flx = Math<R>::floor(sx), fly = Math<R>::floor(sy);
  const int mx = (int)flx, my = (int)fly;
  if (g.pm == EMCGPU_PM_CIC) {
    const R wx = sx - flx, wy = sy - fly;
    R a0x, a0y, a1x, a1y, a2x, a2y, a3x, a3y;
    acc_at(g, acc, mx, my, a0x, a0y);
    acc_at(g, acc, mx + 1, my, a1x, a1y);
    acc_at(g, acc, mx, my + 1, a2x, a2y);
    acc_at(g, acc, mx + 1, my + 1, a3x, a3y);
    fx = a0x * wx * wy + a1x * (R(1) - wx) * wy + a2x * wx * (R(1) - wy) +
         a3x * (R(1) - wy) * (R(1) - wy);
    fy = a0y * wx * wy + a1y * (R(1) - wx) * wy + a2y * wx * (R(1) - wy) +
         a3y * (R(1) - wy) * (R(1) - wy);
  } else {
    R e0x, e0y, e1x, e1y, e2x, e2y;
    acc_at(g, acc, mx, my, e0x, e0y);
    acc_at(g, acc, mx, my + 1, e1x, e1y);
    acc_at(g, acc, mx + 1, my, e2x, e2y);
    fx = (e0x + e1x) / R(2);
    fy = (e0y + e2y) / R(2);
  }
}
template <typename R>
__device__ __forceinline__ void interp(const DevGeom<R> &g, R x, R y, R &fx, R &fy) {
  interp(g, g.acc, x, y, fx, fy);
}

// assignToMesh of one particle (one carrier): add(ix, iy, weight) for the
// scheme's nodes inside the grid (NGP emcNGPScheme.hpp:42-52, CIC
// emcCICScheme.hpp:71-86 with node (floor x, floor y) taking wX * wY, NEC
// emcNECScheme.hpp:62-75)
template <typename R, typename F>
__device__ __forceinline__ void assign(const DevGeom<R> &g, R x, R y, F &&add) {
#pragma clang fp contract(off)
  auto put = [&](int ix, int iy, R w) {
    if ((unsigned)ix < (unsigned)g.nx && (unsigned)iy < (unsigned)g.ny)
      add(iy * g.nx + ix, w);
  };
  if (g.pm == EMCGPU_PM_NGP) {
    put(pos_coord(x, g.hx), pos_coord(y, g.hy), R(1));
    return;
  }
  const R sx = Math<R>::div(x, g.hx), sy = Math<R>::div(y, g.hy);
  const R flx = Math<R>::floor(sx), fly = Math<R>::floor(sy);
  const int mx = (int)flx, my = (int)fly;
  if (g.pm == EMCGPU_PM_CIC) {
    const R wx = sx - flx, wy = sy - fly;
    put(mx, my, wx * wy);
    put(mx + 1, my, (R(1) - wx) * wy);
    put(mx, my + 1, wx * (R(1) - wy));
    put(mx + 1, my + 1, (R(1) - wx) * (R(1) - wy));
  } else {
    put(mx, my, R(0.25));
    put(mx + 1, my, R(0.25));
    put(mx, my + 1, R(0.25));
    put(mx + 1, my + 1, R(0.25));
  }
}

// ---------------------------------------------------------------------------
// free flight with position (emcParticleDrift.hpp:11-35) and the boundary
// (handleParticleAtBoundary, emcParticleDrift.hpp:41-66): returns true if
// the particle left through an ohmic contact (position clamped onto it);
// other boundaries reflect specularly (emcScatterHandler.hpp:177-189)
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ bool drift_dev(const InstDev<R> *P, const DevGeom<R> &g, R t, R ax,
                                          R ay, R &x, R &y, R &kx, R &ky, R &kz, R &E, int v,
                                          int sub) {
  const ValleyDev<R> &V = P->valley[v];
  const R *M = V.hv[sub];
  const R kx0 = kx, ky0 = ky, kz0 = kz;
  // k += M F t / hbar (force has no z component)
  kx += (M[0] * ax + M[1] * ay) * t;
  ky += (M[1] * ax + M[3] * ay) * t;
  kz += (M[2] * ax + M[4] * ay) * t;
  E = energy_of(V, kx, ky, kz);
  // dx = hbar M (k + k0) / 2 t / m_c(E)  [nm]: cv [m/s per 1/nm] * t [fs] * 1e-6
  const R mx = (kx + kx0) * R(0.5), my = (ky + ky0) * R(0.5), mz = (kz + kz0) * R(0.5);
  const R f = Math<R>::div(V.cv * t * R(1e-6), R(1) + R(2) * V.alpha * E);
  x += (M[0] * mx + M[1] * my + M[2] * mz) * f;
  y += (M[1] * mx + M[3] * my + M[4] * mz) * f;
  if (x < R(0) || x > g.maxx || y < R(0) || y > g.maxy) {
    // std::max(0, std::min(pos, maxPos))
    R bx = (g.maxx < x) ? g.maxx : x;
    bx = (R(0) < bx) ? bx : R(0);
    R by = (g.maxy < y) ? g.maxy : y;
    by = (R(0) < by) ? by : R(0);
    const int cx = pos_coord(bx, g.hx), cy = pos_coord(by, g.hy);
    if ((unsigned)cx < (unsigned)g.nx && (unsigned)cy < (unsigned)g.ny &&
        g.cnode[cy * g.nx + cx] >= 0) {
      x = bx;
      y = by;
      return true;
    }
    if (x < R(0)) {
      x = -x;
      kx = -kx;
    } else if (x > g.maxx) {
      x = R(2) * g.maxx - x;
      kx = -kx;
    }
    if (y < R(0)) {
      y = -y;
      ky = -ky;
    } else if (y > g.maxy) {
      y = R(2) * g.maxy - y;
      ky = -ky;
    }
  }
  return false;
}

// x component of getVelocity in the device frame [m/s]
// (emcNonParabolicAnistropValley.hpp:126-136; the parabolic kinds have
// alpha = 0): hbar M k / (m_c sqrt(1 + 4 alpha gamma(E)))
template <typename R>
__device__ __forceinline__ double vx_dev(const InstDev<R> *P, int v, int sub, R kx, R ky, R kz,
                                         R E) {
  const ValleyDev<R> &V = P->valley[v];
  const R *M = V.hv[sub];
  const R g = E * (R(1) + V.alpha * E);
  const R npf = Math<R>::sqrt(R(1) + R(4) * V.alpha * g);
  return (double)Math<R>::div(V.cv * (M[0] * kx + M[1] * ky + M[2] * kz), npf);
}

// one particle, one step of driftScatterParticles
// (emcBasicParticleHandler.hpp:100-131): force at the current position,
// drift min(tau, dt); while time remains and the particle is inside:
// scatter, tau += tau', force again, drift min(rem, tau'), rem -= tau';
// finally tau -= dt
template <typename R, int MODE>
__device__ __forceinline__ bool move_dev(const InstDev<R> *P, const R *__restrict__ cum,
                                         const DevGeom<R> &g, const R *acc, R &x, R &y, R &kx,
                                         R &ky, R &kz, R &E, R &tau, int &v, int &sub,
                                         StepRng<R, MODE> &rng, int &nev) {
  const R dt = P->dt;
  R ax, ay;
  interp(g, acc, x, y, ax, ay);
  bool removed;
  uint32_t e = 0;
  int nreal = 0;
  if (tau >= dt) {
    removed = drift_dev(P, g, dt, ax, ay, x, y, kx, ky, kz, E, v, sub);
  } else {
    removed = drift_dev(P, g, tau, ax, ay, x, y, kx, ky, kz, E, v, sub);
    R rem = dt - tau;
    // every pass draws tau' > 0, so rem strictly decreases; the event cap
    // bounds the loop regardless
    while (rem > R(0) && !removed && e < kMaxEventsPerStep) {
      // the tables of the region of the last drift (driftParticle,
      // emcAbstractParticleHandler.hpp:244-247)
      const InstDev<R> *Pr =
          g.region ? P + g.region[pos_coord(y, g.hy) * g.nx + pos_coord(x, g.hx)] : P;
      const R newTau = scatter_event<R, MODE>(Pr, cum, kx, ky, kz, E, v, sub, e, rng, nreal);
      tau += newTau;
      interp(g, acc, x, y, ax, ay);
      removed = drift_dev(P, g, (newTau < rem) ? newTau : rem, ax, ay, x, y, kx, ky, kz, E, v,
                          sub);
      rem -= newTau;
    }
  }
  tau -= dt;
  nev = (int)e;
  return removed;
}

// a new particle at grid node (cx, cy): initParticlePos
// (emcParticleInitialization.hpp:13-29), the thermal k-space state of
// emcElectron::generateInitialParticle (emcElectron.hpp:75-90) with the
// boundary flips of emcParticleInitialization.hpp:36-51, on the Philox
// stream (step word, pid) of the init domain
template <typename R>
__device__ __forceinline__ void new_particle(const InstDev<R> *P, const DevGeom<R> &g, int cx,
                                             int cy, uint32_t stw, uint32_t pid, uint32_t k0,
                                             uint32_t k1, R &x, R &y, R &kx, R &ky, R &kz, R &E,
                                             R &tau, uint32_t &vs) {
  R u[8];
#pragma unroll
  for (int j = 0; j < 8; ++j)
    u[j] = stream_uniform<R>((uint32_t)j, stw, pid, kDomInit, k0, k1);
  const R lo = R(1e-6), sc = R(1 - 1e-6);
  auto place = [&](int c, int n, R h, R uu) -> R {
    if (c == n - 1)
      return ((R)c - uu * R(0.5)) * h;
    if (c == 0)
      return uu * R(0.5) * h;
    return ((R)c + uu - R(0.5)) * h;
  };
  x = place(cx, g.nx, g.hx, u[0]);
  y = place(cy, g.ny, g.hy, u[1]);
  int v = (int)Math<R>::floor((R)P->n_valleys * (u[2] * sc + lo));
  v = v < P->n_valleys ? v : P->n_valleys - 1;
  const ValleyDev<R> &V = P->valley[v];
  int sub = (int)Math<R>::floor((R)V.deg * (u[3] * sc + lo));
  sub = sub < V.deg ? sub : V.deg - 1;
  E = R(-1.5) * P->vt * Math<R>::log(u[4] * sc + lo);
  random_direction(normk_of(V, E), u[5], u[6], kx, ky, kz);
  if ((cx == 0 && kx < R(0)) || (cx == g.nx - 1 && kx > R(0)))
    kx = -kx;
  if ((cy == 0 && ky < R(0)) || (cy == g.ny - 1 && ky > R(0)))
    ky = -ky;
  // getNewTau with the tau0 of the node's region (emcElectron.hpp:79-87)
  const InstDev<R> *Pr = g.region ? P + g.region[cy * g.nx + cx] : P;
  tau = -Math<R>::log(u[7] * sc + lo) * Pr->tau0[v];
  vs = ((uint32_t)v << 16) | (uint32_t)sub;
}

// ---------------------------------------------------------------------------
// frozen-field drift (emcgpu_device_drift): one thread per particle, all
// steps fused; a particle that leaves stops and is reported
// ---------------------------------------------------------------------------
template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_dev_drift(DevStepArgs<R> a) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n)
    return;
  const InstDev<R> *P = a.inst;
  R x = a.p.x[i], y = a.p.y[i], kx = a.p.kx[i], ky = a.p.ky[i], kz = a.p.kz[i];
  R E = a.p.E[i], tau = a.p.tau[i];
  const uint32_t vs = a.p.vs[i];
  int v = (int)(vs >> 16), sub = (int)(vs & 0xffffu);
  bool out = a.exit_step[i] >= 0 || vs == kHole;
  StepRng<R, MODE> rng;
  if constexpr (MODE == 0) {
    rng.pid = a.p.pid[i];
    rng.inst = kDomMove;
    rng.k0 = a.key0;
    rng.k1 = a.key1;
  } else {
    rng.rs.p = a.draws + i * a.draw_stride;
    rng.rs.len = a.draw_stride;
    rng.rs.pos = (long long)a.cursor[i];
    rng.rs.over = false;
  }
  double grain = (MODE == 1 && a.grain_tau) ? a.grain_tau[i] : 1.0;
  for (int s = 0; s < a.n_steps; ++s) {
    int nev = 0;
    if (!out) {
      if constexpr (MODE == 0)
        rng.step = a.step + (uint32_t)s;
      const bool removed = move_dev<R, MODE>(P, a.cum, a.g, a.g.acc, x, y, kx, ky, kz, E, tau, v, sub,
                                             rng, nev);
      if constexpr (MODE == 1) {
        // grain bookkeeping (emcBasicParticleHandler.hpp:134-138), no grain
        // mechanism: a new grainTau of -ln U[1e-6,1) * 1 s while inside
        grain -= a.dt_s;
        if (grain <= 0 && !removed)
          grain = -::log(rng.rs.uniform(1e-6, 1.)) * 1.0;
      }
      if (removed) {
        out = true;
        a.exit_step[i] = s;
        a.exit_contact[i] = a.g.cnode[pos_coord(y, a.g.hy) * a.g.nx + pos_coord(x, a.g.hx)];
      }
    }
    if (a.tr_pos) {  // SI: m, 1/m
      const long long o = (long long)s * a.n + i;
      a.tr_pos[2 * o] = (double)x * 1e-9;
      a.tr_pos[2 * o + 1] = (double)y * 1e-9;
      a.tr_k[3 * o] = (double)kx * 1e9;
      a.tr_k[3 * o + 1] = (double)ky * 1e9;
      a.tr_k[3 * o + 2] = (double)kz * 1e9;
      a.tr_E[o] = (double)E;
      a.tr_events[o] = nev;
    }
  }
  a.p.x[i] = x;
  a.p.y[i] = y;
  a.p.kx[i] = kx;
  a.p.ky[i] = ky;
  a.p.kz[i] = kz;
  a.p.E[i] = E;
  a.p.tau[i] = tau;
  if (vs != kHole)
    a.p.vs[i] = ((uint32_t)v << 16) | (uint32_t)sub;
  if constexpr (MODE == 1) {
    a.cursor[i] = (uint64_t)rng.rs.pos;
    if (a.grain_tau)
      a.grain_tau[i] = grain;
    if (rng.rs.over)
      atomicAdd(a.draw_overflow, 1);
  }
}

// ---------------------------------------------------------------------------
// the particle step of a self-consistent run
// ---------------------------------------------------------------------------
// particle range of segment seg (one wave each) of an array of n entries
__device__ __forceinline__ void segment_range(long long n, long long n_seg, long long seg,
                                              long long &lo, long long &hi) {
  long long len = (n + n_seg - 1) / n_seg;
  len = (len + 63) & ~63LL;
  lo = seg * len < n ? seg * len : n;
  hi = lo + len < n ? lo + len : n;
}

// The step kernel runs one block per CU on a contiguous range of chunks of
// 64 consecutive particles (one per lane); the block's waves take the
// chunks one at a time from an LDS counter, so waves that progress faster
// take more of them, and load the state of their next chunk while the
// current one moves.  (A copy of the force grid in LDS measured slower: the
// copy costs more than the L2 gathers it saves.)
template <typename R>
__global__ __launch_bounds__(DevStepShape<R>::threads) void k_dev_step(DevStepArgs<R> a) {
  constexpr int NT = DevStepShape<R>::threads, NW = DevStepShape<R>::waves;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ int ex_cnt[kMaxContacts];
  __shared__ double vx_w[NW];
  __shared__ int q_next;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nn = a.g.nn;
  R *grid = (R *)lds_raw;
  if (a.lds_dep)
    for (int k = tid; k < nn; k += NT)
      grid[k] = R(0);
  if (tid < kMaxContacts)
    ex_cnt[tid] = 0;
  if (tid == 0)
    q_next = 0;
  __syncthreads();

  const long long n_total = a.state->n_total;
  const uint32_t stw = (uint32_t)a.state->step;
  const long long n_chunks = (n_total + 63) / 64;
  const long long c_lo = n_chunks * blockIdx.x / gridDim.x;
  const long long c_hi = n_chunks * (blockIdx.x + 1) / gridDim.x;
  const InstDev<R> *P = a.inst;
  double vxs = 0.0;  // channel current: this lane's sum of v_x
  auto take = [&]() -> long long {
    int t = 0;
    if (lane == 0)
      t = atomicAdd(&q_next, 1);
    return c_lo + __shfl(t, 0, 64);
  };
  struct Part {
    R x, y, kx, ky, kz, E, tau;
    uint32_t vs, pid;
  };
  auto load = [&](long long c, Part &q) {
    const long long i = c * 64 + lane;
    q.vs = (c < c_hi && i < n_total) ? a.p.vs[i] : kHole;
    if (q.vs != kHole) {
      q.x = a.p.x[i];
      q.y = a.p.y[i];
      if (a.move) {
        q.kx = a.p.kx[i];
        q.ky = a.p.ky[i];
        q.kz = a.p.kz[i];
        q.E = a.p.E[i];
        q.tau = a.p.tau[i];
        q.pid = a.p.pid[i];
      }
    }
  };
  long long c = take();
  Part nxt;
  load(c, nxt);
  while (c < c_hi) {
    const long long cn = take();
    const Part cur = nxt;
    load(cn, nxt);
    const long long i = c * 64 + lane;
    int slot = -1;
    if (cur.vs != kHole) {
      R x = cur.x, y = cur.y;
      bool removed = false;
      if (a.move) {
        R kx = cur.kx, ky = cur.ky, kz = cur.kz, E = cur.E, tau = cur.tau;
        int v = (int)(cur.vs >> 16), sub = (int)(cur.vs & 0xffffu);
        StepRng<R, 0> rng;
        rng.step = stw;
        rng.pid = cur.pid;
        rng.inst = kDomMove;
        rng.k0 = a.key0;
        rng.k1 = a.key1;
        int nev;
        removed =
            move_dev<R, 0>(P, a.cum, a.g, a.g.acc, x, y, kx, ky, kz, E, tau, v, sub, rng, nev);
        if (removed) {
          a.p.vs[i] = kHole;
          const int cidx = a.g.cnode[pos_coord(y, a.g.hy) * a.g.nx + pos_coord(x, a.g.hx)];
          atomicAdd(&ex_cnt[cidx], 1);
        } else {
          a.p.x[i] = x;
          a.p.y[i] = y;
          a.p.kx[i] = kx;
          a.p.ky[i] = ky;
          a.p.kz[i] = kz;
          a.p.E[i] = E;
          a.p.tau[i] = tau;
          a.p.vs[i] = ((uint32_t)v << 16) | (uint32_t)sub;
          if (a.chan && x >= a.chan_x0 && x <= a.chan_x1)
            vxs += vx_dev(P, v, sub, kx, ky, kz, E);
        }
      }
      if (!removed) {
        if (a.lds_dep)
          assign(a.g, x, y, [&](int k, R w) { atomicAdd(&grid[k], w); });
        else
          assign(a.g, x, y, [&](int k, R w) { atomicAdd(&a.dep[k], w); });
        const int cx = pos_coord(x, a.g.hx), cy = pos_coord(y, a.g.hy);
        if ((unsigned)cx < (unsigned)a.g.nx && (unsigned)cy < (unsigned)a.g.ny)
          slot = a.g.cslot[cy * a.g.nx + cx];
      }
    }
    if (i < n_total)
      a.pslot[i] = (int16_t)slot;
    c = cn;
  }
  if (a.chan) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      vxs += __shfl_xor(vxs, off, 64);
    if (lane == 0)
      vx_w[wave] = vxs;
  }
  __syncthreads();
  if (a.chan && tid == 0) {
    double b = 0.0;
    for (int w = 0; w < NW; ++w)
      b += vx_w[w];
    a.chan_part[blockIdx.x] = b;
  }
  if (a.lds_dep)
    for (int k = tid; k < nn; k += NT)
      a.dep[(long long)blockIdx.x * nn + k] = grid[k];
  if (tid < a.n_contacts && a.exit_part)
    a.exit_part[blockIdx.x * a.n_contacts + tid] = ex_cnt[tid];
}

// rank of each particle at an ohmic node among the particles of its segment
// at that node, in array order (one wave per segment), from the slots the
// step kernel recorded; the staged entries and per-slot counts feed the
// contact scan
template <typename R>
__device__ __forceinline__ void dev_rank(const DevStepArgs<R> &a, long long n_seg, long long seg,
                                         uint32_t *mycnt) {
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < a.n_slots; k += 64)
    mycnt[k] = 0u;
  const long long n_total = a.state->n_total;
  long long lo, hi;
  segment_range(n_total, n_seg, seg, lo, hi);
  const uint64_t lt = (1ull << lane) - 1ull;
  uint32_t nst = 0;
  for (long long base = lo; base < hi; base += 256) {
    int sl[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = base + 64 * u + lane;
      sl[u] = i < hi ? (int)a.pslot[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = base + 64 * u + lane;
      const int slot = sl[u];
      const uint64_t at = __ballot(slot >= 0);
      if (!at)
        continue;
      // one pass per distinct slot of the 64 entries, registers only: the
      // lane's rank among the lanes with its slot, and the group size at
      // the group's lowest lane
      uint64_t pend = at;
      uint32_t lrank = 0, gsize = 0;
      while (pend) {
        const int leader = __ffsll((unsigned long long)pend) - 1;
        const int s = __builtin_amdgcn_readlane(slot, leader);
        const uint64_t m = __ballot(slot == s);
        if (slot == s)
          lrank = (uint32_t)__popcll(m & lt);
        if (lane == leader)
          gsize = (uint32_t)__popcll(m);
        pend &= ~m;
      }
      // then the segment's counts: every lane reads its slot's count, the
      // groups' lowest lanes advance them
      const uint32_t prior = slot >= 0 ? mycnt[slot] : 0u;
      if (gsize)
        mycnt[slot] = prior + gsize;
      if (slot >= 0)
        a.stage[lo + nst + (uint32_t)__popcll(at & lt)] =
            make_uint2((uint32_t)i, ((uint32_t)slot << 22) | (prior + lrank));
      nst += (uint32_t)__popcll(at);
    }
  }
  // slot-major, so the contact scan of a slot reads its segments' counts
  // contiguously
  for (int k = lane; k < a.n_slots; k += 64)
    a.seg_cnt[(long long)k * n_seg + seg] = mycnt[k];
  if (lane == 0)
    a.seg_nst[seg] = nst;
}

// first pass of the FP64 sum of the block grids: rows [b0, b1) of node k
template <typename R>
__device__ __forceinline__ void dev_reduce_rows(const DevContactArgs<R> &a, int k, int chunk) {
  const int nn = a.g.nn;
  if (k >= nn)
    return;
  const int b0 = chunk * a.rows_per_chunk;
  const int b1 = b0 + a.rows_per_chunk < a.n_blocks_dep ? b0 + a.rows_per_chunk : a.n_blocks_dep;
  double s = 0.0;
  for (int b = b0; b < b1; ++b)
    s += (double)a.dep[(long long)b * nn + k];
  a.dpart[(long long)chunk * nn + k] = s;
  if (a.n_blocks_dep == 1)  // the atomic deposit accumulates in row 0
    a.dep[k] = R(0);
}

// after the step kernel, two independent jobs in one launch: blocks
// [0, n_seg / 8) rank the contact particles (one wave per segment), the
// others sum the block grids in chunks of rows (one node per thread)
template <typename R>
__global__ __launch_bounds__(kDevStepThreads) void k_dev_rank_rows(DevStepArgs<R> a,
                                                                   DevContactArgs<R> ca,
                                                                   int n_rank_blocks) {
  extern __shared__ uint32_t rank_cnt[];  // [wave][slot]
  const int tid = threadIdx.x, wave = tid >> 6;
  if ((int)blockIdx.x < n_rank_blocks) {
    dev_rank(a, (long long)n_rank_blocks * kDevStepWaves,
             (long long)blockIdx.x * kDevStepWaves + wave, rank_cnt + wave * a.n_slots);
    return;
  }
  const int r = (int)blockIdx.x - n_rank_blocks;
  const int node_blocks = (ca.g.nn + kDevStepThreads - 1) / kDevStepThreads;
  dev_reduce_rows(ca, (r % node_blocks) * kDevStepThreads + tid, r / node_blocks);
}

// ---------------------------------------------------------------------------
// ohmic contacts: per node, array-order offsets of the segments' particles,
// then keep while nrPart < expected, inject ceil(expected - kept)
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_dev_contact_scan(DevContactArgs<R> a) {
  __shared__ uint32_t sc[256];
  const int slot = blockIdx.x, tid = threadIdx.x;
  const int S = a.n_segments;
  const uint32_t *cnt = a.seg_cnt + (long long)slot * S;  // row of the [slot][segment] counts
  uint32_t *off = a.seg_off + (long long)slot * S;
  const int per = (S + 255) / 256;
  const int b0 = tid * per < S ? tid * per : S;
  const int b1 = b0 + per < S ? b0 + per : S;
  // the thread's counts, sixteen loads in flight at a time
  uint32_t tot = 0;
  for (int b = b0; b < b1; b += 16) {
    uint32_t v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u)
      v[u] = b + u < b1 ? cnt[b + u] : 0u;
#pragma unroll
    for (int u = 0; u < 16; ++u)
      tot += v[u];
  }
  sc[tid] = tot;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const uint32_t t = tid >= d ? sc[tid - d] : 0u;
    __syncthreads();
    sc[tid] += t;
    __syncthreads();
  }
  uint32_t run = sc[tid] - tot;
  for (int b = b0; b < b1; b += 16) {
    uint32_t v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u)
      v[u] = b + u < b1 ? cnt[b + u] : 0u;
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (b + u < b1) {
        off[b + u] = run;
        run += v[u];
      }
  }
  // the node's counts (no atomics: the resolve kernel totals them per
  // contact)
  if (tid == 0) {
    const long long total = sc[255];
    const double e = a.slot_exp[slot];
    const long long ce = (long long)ceil(e);
    const long long kept = total < ce ? total : ce;
    const double diff = e - (double)kept;
    const long long inj = diff > 0.0 ? (long long)ceil(diff) : 0;
    a.slot_total[slot] = (int)total;
    a.slot_kept[slot] = (int)kept;
    a.slot_inj[slot] = (int)inj;
  }
}

// deletions of the excess at ohmic nodes (one wave per segment) and
// injections (one thread per new particle, after the deletion blocks)
template <typename R>
__global__ __launch_bounds__(256) void k_dev_resolve(DevContactArgs<R> a) {
  __shared__ int pre[kMaxSlots + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long n_total = a.state->n_total;
  const int nA = (a.n_segments + 3) / 4;
  if ((int)blockIdx.x < nA) {
    const long long seg = (long long)blockIdx.x * 4 + wave;
    if (seg >= a.n_segments)
      return;
    long long lo, hi;
    segment_range(n_total, a.n_segments, seg, lo, hi);
    const uint32_t nst = a.seg_nst[seg];
    for (uint32_t j = lane; j < nst; j += 64) {
      const uint2 e = a.stage[lo + j];
      const int slot = (int)(e.y >> 22);
      const uint32_t rank = a.seg_off[(long long)slot * a.n_segments + seg] + (e.y & 0x3fffffu);
      if (rank >= (uint32_t)a.slot_kept[slot]) {
        const long long i = e.x;
        const R x = a.p.x[i];
        if (a.chan && x >= a.chan_x0 && x <= a.chan_x1) {
          const uint32_t vs = a.p.vs[i];
          atomicAdd(a.chan_corr, -vx_dev(a.inst, (int)(vs >> 16), (int)(vs & 0xffffu), a.p.kx[i],
                                         a.p.ky[i], a.p.kz[i], a.p.E[i]));
        }
        a.p.vs[i] = kHole;
        assign(a.g, x, a.p.y[i], [&](int k, R w) { atomicAdd(&a.corr[k], -(double)w); });
      }
    }
    return;
  }
  // injections
  const long long bi = (long long)blockIdx.x - nA;
  const long long t = bi * 256 + tid;
  {  // prefix of the injections over the ohmic nodes (4 nodes per thread)
    __shared__ int ws[256];
    int v[4], run = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = 4 * tid + u;
      v[u] = q < a.n_slots ? a.slot_inj[q] : 0;
      run += v[u];
    }
    ws[tid] = run;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int o = tid >= d ? ws[tid - d] : 0;
      __syncthreads();
      ws[tid] += o;
      __syncthreads();
    }
    int acc = ws[tid] - run;
    if (tid == 0)
      pre[0] = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = 4 * tid + u;
      acc += v[u];
      if (q < a.n_slots)
        pre[q + 1] = acc;
    }
  }
  __shared__ int per_contact[3][kMaxContacts];
  if (bi == 0 && tid < 3 * kMaxContacts)
    per_contact[tid / kMaxContacts][tid % kMaxContacts] = 0;
  __syncthreads();
  long long T = pre[a.n_slots];
  const long long room = a.state->capacity - n_total;
  if (T > room)
    T = room;
  if (bi == 0) {
    // the step's records: deletions and injections per contact from the
    // nodes' counts, exits from the step blocks' counts
    for (int q = tid; q < a.n_slots; q += 256) {
      const int c = a.slot_contact[q];
      atomicAdd(&per_contact[0][c], a.slot_total[q] - a.slot_kept[q]);
      atomicAdd(&per_contact[1][c], a.slot_inj[q]);
    }
    if (a.exit_part)
      for (int e = tid; e < a.n_step_blocks * a.n_contacts; e += 256)
        atomicAdd(&per_contact[2][e % a.n_contacts], a.exit_part[e]);
    __syncthreads();
    if (tid == 0) {
      if (T < pre[a.n_slots])
        a.state->overflow = 1;
      a.state->inj_total = T;
    }
    const long long row = a.state->step - a.state->hist_base;
    if (tid < a.n_contacts) {
      const long long o = row * a.n_contacts + tid;
      if (a.hist.dels)
        a.hist.dels[o] += per_contact[0][tid];
      if (a.hist.injs)
        a.hist.injs[o] += per_contact[1][tid];
      if (a.hist.exits && a.exit_part)
        a.hist.exits[o] += per_contact[2][tid];
    }
  }
  if (t >= T)
    return;
  int lo = 0, hi = a.n_slots;  // pre[lo] <= t < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  const int node = a.slot_node[lo];
  const int cx = node % a.g.nx, cy = node / a.g.nx;
  const long long idx = n_total + t;
  const uint32_t pid = (uint32_t)(a.state->next_pid + t);
  R x, y, kx, ky, kz, E, tau;
  uint32_t vs;
  new_particle(a.inst, a.g, cx, cy, (uint32_t)a.state->step, pid, a.key0, a.key1, x, y, kx, ky,
               kz, E, tau, vs);
  a.p.x[idx] = x;
  a.p.y[idx] = y;
  a.p.kx[idx] = kx;
  a.p.ky[idx] = ky;
  a.p.kz[idx] = kz;
  a.p.E[idx] = E;
  a.p.tau[idx] = tau;
  a.p.vs[idx] = vs;
  a.p.pid[idx] = pid;
  assign(a.g, x, y, [&](int k, R w) { atomicAdd(&a.corr[k], (double)w); });
  if (a.chan && x >= a.chan_x0 && x <= a.chan_x1)
    atomicAdd(a.chan_corr, vx_dev(a.inst, (int)(vs >> 16), (int)(vs & 0xffffu), kx, ky, kz, E));
}

// carriers per node = FP64 sum of the block grids (fixed order) + the
// corrections; concentration (updateCurrentParticleConcentrations,
// emcSimulationResults.hpp:98-116); commit of the step's counters
// first pass: chunk c of the block grids' rows summed per node (FP64)
template <typename R>
__global__ __launch_bounds__(256) void k_dev_reduce(DevContactArgs<R> a, int commit) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int nn = a.g.nn;
  if (k < nn) {
    double s = 0.0;
    for (int c = 0; c < a.n_chunks; ++c)
      s += a.dpart[(long long)c * nn + k];
    s += a.corr[k];
    a.corr[k] = 0.0;
    a.nrpart[k] = s;
    const double cn = s * a.cfac[k];
    a.conc[k] = cn;
    if (a.accumulate) {  // updateAverageCharacteristics (emcSimulationResults.hpp:85-93)
      a.avg_pot[k] += a.pot[k];
      a.avg_conc[k] += cn;
    }
  }
  if (commit && a.chan && blockIdx.x == 0 && threadIdx.x < 64) {
    // the step's channel current: block partials in a fixed order, then the
    // corrections of the deleted / injected particles
    const long long row = a.state->step - a.state->hist_base;
    double v = 0.0;
    for (int b = threadIdx.x; b < a.n_chan_blocks; b += 64)
      v += a.chan_part[b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      v += __shfl_xor(v, off, 64);
    if (threadIdx.x == 0) {
      a.hist.chan[row] = (v + *a.chan_corr) * a.chan_scale;
      *a.chan_corr = 0.0;
    }
  }
  if (commit && blockIdx.x == 0 && threadIdx.x == 0) {
    DevState *st = a.state;
    const long long row = st->step - st->hist_base;
    if (a.accumulate)
      st->avg_n += 1;
    long long d = 0;
    for (int c = 0; c < a.n_contacts; ++c) {
      const long long o = row * a.n_contacts + c;
      d += (long long)a.hist.injs[o] - a.hist.exits[o] - a.hist.dels[o];
    }
    st->n_alive += d;
    st->n_total += st->inj_total;
    st->next_pid += st->inj_total;
    st->inj_total = 0;
    if (a.hist.n_part)
      a.hist.n_part[row] = st->n_alive;
    st->step += 1;
  }
}

// zeroes a global grid row (deposits of the atomic path, corrections)
__global__ void k_dev_zero(double *p, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    p[i] = 0.0;
}
template <typename R>
__global__ void k_dev_zero_r(R *p, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    p[i] = R(0);
}
__global__ void k_dev_zero_i(int *p, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    p[i] = 0;
}

// initial particles: particle t at the node whose prefix range holds t
template <typename R>
__global__ __launch_bounds__(256) void k_dev_init(DevContactArgs<R> a, const long long *prefix,
                                                  long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n)
    return;
  int lo = 0, hi = a.g.nn;  // prefix[lo] <= t < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  R x, y, kx, ky, kz, E, tau;
  uint32_t vs;
  new_particle(a.inst, a.g, lo % a.g.nx, lo / a.g.nx, 0xFFFFFFFFu, (uint32_t)t, a.key0, a.key1, x,
               y, kx, ky, kz, E, tau, vs);
  a.p.x[t] = x;
  a.p.y[t] = y;
  a.p.kx[t] = kx;
  a.p.ky[t] = ky;
  a.p.kz[t] = kz;
  a.p.E[t] = E;
  a.p.tau[t] = tau;
  a.p.vs[t] = vs;
  a.p.pid[t] = (uint32_t)t;
}

// ---------------------------------------------------------------------------
// stable compaction of the ensemble (holes out, array order kept)
// ---------------------------------------------------------------------------
constexpr int kCompactTile = 1024;

__global__ __launch_bounds__(256) void k_dev_count_alive(const uint32_t *vs, const DevState *st,
                                                         int *blk) {
  __shared__ int ws[4];
  const long long n = st->n_total;
  const long long i0 = (long long)blockIdx.x * kCompactTile;
  int c = 0;
  for (int j = threadIdx.x; j < kCompactTile; j += 256) {
    const long long i = i0 + j;
    c += (i < n && vs[i] != kHole) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1)
    c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0)
    ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0)
    blk[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// exclusive scan of the block counts in one workgroup
__global__ __launch_bounds__(1024) void k_dev_scan_blocks(int *blk, int nb, DevState *st) {
  __shared__ long long sc[1024];
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = tid * per < nb ? tid * per : nb;
  const int b1 = b0 + per < nb ? b0 + per : nb;
  long long tot = 0;
  for (int b = b0; b < b1; ++b)
    tot += blk[b];
  sc[tid] = tot;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const long long t = tid >= d ? sc[tid - d] : 0;
    __syncthreads();
    sc[tid] += t;
    __syncthreads();
  }
  long long run = sc[tid] - tot;
  for (int b = b0; b < b1; ++b) {
    const int c = blk[b];
    blk[b] = (int)run;
    run += c;
  }
  if (tid == 1023)
    st->n_alive = sc[1023];
}

template <typename R>
__global__ __launch_bounds__(256) void k_dev_compact(DevParticles<R> src, DevParticles<R> dst,
                                                     const int *blk, DevState *st) {
  __shared__ int ws[4];
  const long long n = st->n_total;
  const long long i0 = (long long)blockIdx.x * kCompactTile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int base = blk[blockIdx.x];
  for (int j0 = 0; j0 < kCompactTile; j0 += 256) {
    const long long i = i0 + j0 + threadIdx.x;
    const bool keep = i < n && src.vs[i] != kHole;
    const uint64_t m = __ballot(keep);
    if (lane == 0)
      ws[wave] = __popcll(m);
    __syncthreads();
    int wo = 0;
    for (int w = 0; w < wave; ++w)
      wo += ws[w];
    const int tot = ws[0] + ws[1] + ws[2] + ws[3];
    if (keep) {
      const long long o = base + wo + __popcll(m & ((1ull << lane) - 1ull));
      dst.x[o] = src.x[i];
      dst.y[o] = src.y[i];
      dst.kx[o] = src.kx[i];
      dst.ky[o] = src.ky[i];
      dst.kz[o] = src.kz[i];
      dst.E[o] = src.E[i];
      dst.tau[o] = src.tau[i];
      dst.vs[o] = src.vs[i];
      dst.pid[o] = src.pid[i];
    }
    base += tot;
    __syncthreads();
  }
}

__global__ void k_dev_compact_commit(DevState *st) { st->n_total = st->n_alive; }

// history rows of the next steps start at the current step
__global__ void k_dev_mark(DevState *st) { st->hist_base = st->step; }

// ---------------------------------------------------------------------------
// known-answer entry points on caller positions (metres, FP64)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dev_assign_pos(DevGeom<double> g, const double *pos,
                                                        long long n, double *grid) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n)
    return;
  assign(g, pos[2 * i], pos[2 * i + 1], [&](int k, double w) { atomicAdd(&grid[k], w); });
}

__global__ __launch_bounds__(256) void k_dev_interp_pos(DevGeom<double> g, const double *pos,
                                                        long long n, double charge,
                                                        double *force) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n)
    return;
  double fx, fy;
  interp(g, pos[2 * i], pos[2 * i + 1], fx, fy);
  force[2 * i] = charge * fx;
  force[2 * i + 1] = charge * fy;
}

// ---------------------------------------------------------------------------
// 3-D particle-mesh operators on caller positions and grids (FP64): the
// reference's 3-D NGP and CIC schemes and the grid-point E-field
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pm3_in(const PMGrid3 &g, long long i, long long j, long long k) {
  return i >= 0 && j >= 0 && k >= 0 && i < g.nx && j < g.ny && k < g.nz;
}
__device__ __forceinline__ long long pm3_idx(const PMGrid3 &g, long long i, long long j,
                                             long long k) {
  return i + (long long)g.nx * (j + (long long)g.ny * k);
}
// a node coordinate from round / floor of pos / h; -1 for values no grid
// holds (negative, huge or not a number), whose nodes are then skipped
__device__ __forceinline__ long long pm3_coord(double c) {
  return (c >= 0.0 && c < 1073741824.0) ? (long long)c : -1;
}

// assignToMesh of one carrier per position: NGP puts it on the node at
// round(pos / h) (emcNGPScheme.hpp:36-48); CIC gives the node at the floor
// corner wX wY wZ and its neighbours the complements, each weight a left-
// to-right product (emcCICScheme.hpp:97-120).  Nodes outside the grid are
// skipped.  The grid sums in atomic (unordered) order.
__global__ __launch_bounds__(256) void k_pm3_assign(PMGrid3 g, const double *pos, long long n,
                                                    double *grid) {
#pragma clang fp contract(off)
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= n)
    return;
  const double sx = pos[3 * q] / g.hx, sy = pos[3 * q + 1] / g.hy, sz = pos[3 * q + 2] / g.hz;
  auto put = [&](long long i, long long j, long long k, double w) {
    if (pm3_in(g, i, j, k))
      atomicAdd(&grid[pm3_idx(g, i, j, k)], w);
  };
  if (g.pm == EMCGPU_PM_NGP) {
    put(pm3_coord(round(sx)), pm3_coord(round(sy)), pm3_coord(round(sz)), 1.0);
    return;
  }
  const double fx = floor(sx), fy = floor(sy), fz = floor(sz);
  const long long i = pm3_coord(fx), j = pm3_coord(fy), k = pm3_coord(fz);
  if (i < 0 || j < 0 || k < 0)
    return;
  const double wx = sx - fx, wy = sy - fy, wz = sz - fz;
  put(i, j, k, wx * wy * wz);
  put(i + 1, j, k, (1.0 - wx) * wy * wz);
  put(i, j + 1, k, wx * (1.0 - wy) * wz);
  put(i + 1, j + 1, k, (1.0 - wx) * (1.0 - wy) * wz);
  put(i, j, k + 1, wx * wy * (1.0 - wz));
  put(i + 1, j, k + 1, (1.0 - wx) * wy * (1.0 - wz));
  put(i, j + 1, k + 1, wx * (1.0 - wy) * (1.0 - wz));
  put(i + 1, j + 1, k + 1, (1.0 - wx) * (1.0 - wy) * (1.0 - wz));
}

// interpolateForce: NGP the nearest node's field times the charge
// (emcNGPScheme.hpp:51-66); CIC the eight-node sum with the reference's
// (1 - wY)(1 - wY) weight on the (x+1, y+1) nodes of both z layers, times
// the charge (emcCICScheme.hpp:143-173).  Nodes outside the grid count zero.
__global__ __launch_bounds__(256) void k_pm3_interp(PMGrid3 g, const double *ex,
                                                    const double *ey, const double *ez,
                                                    const double *pos, long long n, double charge,
                                                    double *force) {
#pragma clang fp contract(off)
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= n)
    return;
  const double *E[3] = {ex, ey, ez};
  const double sx = pos[3 * q] / g.hx, sy = pos[3 * q + 1] / g.hy, sz = pos[3 * q + 2] / g.hz;
  auto at = [&](const double *e, long long i, long long j, long long k) {
    return pm3_in(g, i, j, k) ? e[pm3_idx(g, i, j, k)] : 0.0;
  };
  if (g.pm == EMCGPU_PM_NGP) {
    const long long i = pm3_coord(round(sx)), j = pm3_coord(round(sy)), k = pm3_coord(round(sz));
    for (int d = 0; d < 3; ++d)
      force[3 * q + d] = charge * at(E[d], i, j, k);
    return;
  }
  const double fx = floor(sx), fy = floor(sy), fz = floor(sz);
  const long long i = pm3_coord(fx), j = pm3_coord(fy), k = pm3_coord(fz);
  const double wx = sx - fx, wy = sy - fy, wz = sz - fz;
  const bool ok = i >= 0 && j >= 0 && k >= 0;
  for (int d = 0; d < 3; ++d) {
    const double *e = E[d];
    double f = 0.0;
    if (ok)
      f = (at(e, i, j, k) * wx * wy * wz + at(e, i + 1, j, k) * (1.0 - wx) * wy * wz +
           at(e, i, j + 1, k) * wx * (1.0 - wy) * wz +
           at(e, i + 1, j + 1, k) * (1.0 - wy) * (1.0 - wy) * wz +
           at(e, i, j, k + 1) * wx * wy * (1.0 - wz) +
           at(e, i + 1, j, k + 1) * (1.0 - wx) * wy * (1.0 - wz) +
           at(e, i, j + 1, k + 1) * wx * (1.0 - wy) * (1.0 - wz) +
           at(e, i + 1, j + 1, k + 1) * (1.0 - wy) * (1.0 - wy) * (1.0 - wz)) *
          charge;
    force[3 * q + d] = f;
  }
}

// calcEFieldAtGridPts (emcEFieldCalculation.hpp:13-32): off the faces of
// component d, E_d = (phi_prev - phi_next) * vt / (2 h_d); one node per
// thread
__global__ __launch_bounds__(256) void k_pm3_efield(PMGrid3 g, double vt, const double *pot,
                                                    double *ex, double *ey, double *ez) {
#pragma clang fp contract(off)
  const long long m = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nn = (long long)g.nx * g.ny * g.nz;
  if (m >= nn)
    return;
  const long long nxy = (long long)g.nx * g.ny;
  const long long k = m / nxy, r = m - k * nxy, j = r / g.nx, i = r - j * g.nx;
  const long long c[3] = {i, j, k}, ext[3] = {g.nx, g.ny, g.nz}, s[3] = {1, g.nx, nxy};
  const double h[3] = {g.hx, g.hy, g.hz};
  double *E[3] = {ex, ey, ez};
  for (int d = 0; d < 3; ++d)
    if (c[d] != 0 && c[d] != ext[d] - 1)
      E[d][m] = (pot[m - s[d]] - pot[m + s[d]]) * vt / (2.0 * h[d]);
}

// setEFieldBoundaryValues (emcEFieldCalculation.hpp:57-81) after the
// interior pass: on a face normal to d, bc[d][node] -1 (artificial) sets
// E_d = 0, a contact (>= 0) the inner neighbour's value along d
__global__ __launch_bounds__(256) void k_pm3_faces(PMGrid3 g, const signed char *bc, double *ex,
                                                   double *ey, double *ez) {
  const long long m = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nn = (long long)g.nx * g.ny * g.nz;
  if (m >= nn)
    return;
  const long long nxy = (long long)g.nx * g.ny;
  const long long k = m / nxy, r = m - k * nxy, j = r / g.nx, i = r - j * g.nx;
  const long long c[3] = {i, j, k}, s[3] = {1, g.nx, nxy};
  double *E[3] = {ex, ey, ez};
  for (int d = 0; d < 3; ++d) {
    const int b = bc[(long long)d * nn + m];
    if (b == -2)
      continue;
    E[d][m] = b == -1 ? 0.0 : (c[d] == 0 ? E[d][m + s[d]] : E[d][m - s[d]]);
  }
}

// ---------------------------------------------------------------------------
// Poisson (Newton + FP32 multigrid) and E-field on one workgroup
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_max(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < kDevThreads / 64; ++w)
    m = red[w] > m ? red[w] : m;
  return m;
}

__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kDevThreads / 64; ++w)
    s += red[w];
  return s;
}

// ---------------------------------------------------------------------------
// multigrid V-cycle (FP32, LDS) as the preconditioner of an FP64 conjugate
// gradient solve of the Newton system
// ---------------------------------------------------------------------------

// wave-synchronous step: the wave's LDS writes are visible to its lanes
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct BlockTeam {
  static constexpr int size = kDevThreads;
  __device__ __forceinline__ int rank() const { return (int)threadIdx.x; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};
struct WaveTeam {  // wave 0 alone
  static constexpr int size = 64;
  __device__ __forceinline__ int rank() const { return (int)threadIdx.x; }
  __device__ __forceinline__ void sync() const { wave_sync(); }
};

// Pointers of the Poisson kernel come from its argument block in memory, so
// the compiler cannot tell their address space and emits flat accesses,
// which also count against the LDS wait counter; these views are global.
// (the host compilation pass of this file sees no address spaces)
#if defined(__HIP_DEVICE_COMPILE__)
#define EMC_GLOBAL __attribute__((address_space(1)))
#else
#define EMC_GLOBAL
#endif
template <typename T>
__device__ __forceinline__ EMC_GLOBAL T *glb(T *p) {
  return (EMC_GLOBAL T *)p;
}

// k / n for 0 <= k < 2^22, n >= 1, exact: a reciprocal estimate (error below
// one) and one correction.  The index maths of the Poisson passes is the
// bulk of their vector-ALU work; a run-time integer division costs ~30
// instructions.
__device__ __forceinline__ int idiv(int k, int n) {
  int q = (int)((float)k * __builtin_amdgcn_rcpf((float)n));
  const int r = k - q * n;
  q += r >= n ? 1 : 0;
  q -= r < 0 ? 1 : 0;
  return q;
}

// the coarse levels' data in LDS (level l >= 1 at offset lev[l].off - n0)
struct MGShared {
  float *d;        // corrections of every level (level-stacked)
  float *res;      // level-0 residual (restriction input), level-0 shift scratch
  float *rc, *sc;  // right-hand sides and diagonal shifts of the coarse levels
  const DimTab *tab;
  const uint8_t *cmask;
  int n0;
  float hfx, hfy;
};

// diagonal shift of a Dirichlet node: its correction vanishes in a sweep
// without a test
constexpr float kBigShift = 1e30f;

// one red-black Gauss-Seidel colour on coarse level l of A d = r
template <typename R, typename T>
__device__ __forceinline__ void mg_sweep_c(const DevFieldArgs<R> &a, const MGShared &m, int l,
                                           int c, T team) {
  const MGLevel L = a.lev[l];
  float *d = m.d + L.off;
  const float *r = m.rc + (L.off - m.n0);
  const float *sg = m.sc + (L.off - m.n0);  // kBigShift at Dirichlet nodes
  const DimTab *TX = m.tab + L.tx, *TY = m.tab + L.ty;
  const int hr = (L.nx + 1) >> 1;
  for (int t = team.rank(); t < hr * L.ny; t += T::size) {
    const int j = idiv(t, hr);
    const int i = 2 * (t - j * hr) + ((j + c) & 1);
    if (i >= L.nx)
      continue;
    const int k = j * L.nx + i;
    const DimTab tx = TX[i], ty = TY[j];
    const float cl = m.hfx * ty.w * tx.inv_l, cr = m.hfx * ty.w * tx.inv_r;
    const float cd = m.hfy * tx.w * ty.inv_l, cu = m.hfy * tx.w * ty.inv_r;
    const int il = i > 0 ? k - 1 : k + 1, ir = i < L.nx - 1 ? k + 1 : k - 1;
    const int jd = j > 0 ? k - L.nx : k + L.nx, ju = j < L.ny - 1 ? k + L.nx : k - L.nx;
    // the smoother is a preconditioner: the hardware reciprocal suffices
    d[k] = (r[k] + cl * d[il] + cr * d[ir] + cd * d[jd] + cu * d[ju]) *
           __builtin_amdgcn_rcpf(cl + cr + cd + cu + sg[k]);
  }
  team.sync();
}

// full weighting (the transpose of bilinear prolongation, mirror images
// folded in) of a fine-level field src (level l, dense nx * ny) into the
// coarse level l+1 array dst; masked coarse nodes get 0; zero_d clears the
// coarse correction
template <typename R, typename T>
__device__ __forceinline__ void mg_restrict_to(const DevFieldArgs<R> &a, const MGShared &m, int l,
                                               const float *src, float *dst, bool zero_d,
                                               T team, bool keep_masked = false) {
  const MGLevel L = a.lev[l], C = a.lev[l + 1];
  const uint8_t *cmask = m.cmask + (C.off - m.n0);
  const DimTab *CX = m.tab + C.tx, *CY = m.tab + C.ty;
  float *dc = m.d + C.off;
  for (int K = team.rank(); K < C.nx * C.ny; K += T::size) {
    if (zero_d)
      dc[K] = 0.0f;
    if (cmask[K]) {
      if (!keep_masked)
        dst[K] = 0.0f;
      continue;
    }
    const int J = idiv(K, C.nx), I = K - J * C.nx;
    const DimTab cx = CX[I], cy = CY[J];
    const float wx[3] = {cx.wl, 1.0f, cx.wr}, wy[3] = {cy.wl, 1.0f, cy.wr};
    float s = 0.0f;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if (wy[b] == 0.0f)
        continue;
      const int jj = cy.f + b - 1;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if (wx[q] == 0.0f)
          continue;
        s += wx[q] * wy[b] * src[jj * L.nx + cx.f + q - 1];
      }
    }
    dst[K] = s;
  }
  team.sync();
}

// residual of coarse level l into res (dense), then its restriction
template <typename R, typename T>
__device__ __forceinline__ void mg_residual_c(const DevFieldArgs<R> &a, const MGShared &m, int l,
                                              float *res, T team) {
  const MGLevel L = a.lev[l];
  const float *d = m.d + L.off;
  const float *r = m.rc + (L.off - m.n0);
  const float *sg = m.sc + (L.off - m.n0);
  const uint8_t *mask = m.cmask + (L.off - m.n0);
  const DimTab *TX = m.tab + L.tx, *TY = m.tab + L.ty;
  for (int k = team.rank(); k < L.nx * L.ny; k += T::size) {
    if (mask[k]) {
      res[k] = 0.0f;
      continue;
    }
    const int j = idiv(k, L.nx), i = k - j * L.nx;
    const DimTab tx = TX[i], ty = TY[j];
    const float cl = m.hfx * ty.w * tx.inv_l, cr = m.hfx * ty.w * tx.inv_r;
    const float cd = m.hfy * tx.w * ty.inv_l, cu = m.hfy * tx.w * ty.inv_r;
    const int il = i > 0 ? k - 1 : k + 1, ir = i < L.nx - 1 ? k + 1 : k - 1;
    const int jd = j > 0 ? k - L.nx : k + L.nx, ju = j < L.ny - 1 ? k + L.nx : k - L.nx;
    res[k] = r[k] - ((cl + cr + cd + cu + sg[k]) * d[k] -
                     (cl * d[il] + cr * d[ir] + cd * d[jd] + cu * d[ju]));
  }
  team.sync();
  mg_restrict_to(a, m, l, res, m.rc + (a.lev[l + 1].off - m.n0), true, team);
}

// d_l += P d_{l+1} on coarse level l >= 0 (bilinear in the fine positions);
// level 0 skips its Dirichlet nodes through dmask
template <typename R, typename T>
__device__ __forceinline__ void mg_prolong(const DevFieldArgs<R> &a, const MGShared &m, int l,
                                           T team) {
  const MGLevel L = a.lev[l], C = a.lev[l + 1];
  float *d = m.d + L.off;
  const float *dc = m.d + C.off;
  const uint8_t *mask = l == 0 ? a.dmask : m.cmask + (L.off - m.n0);
  const DimTab *TX = m.tab + L.tx, *TY = m.tab + L.ty;
  for (int k = team.rank(); k < L.nx * L.ny; k += T::size) {
    if (mask[k])
      continue;
    const int j = idiv(k, L.nx), i = k - j * L.nx;
    const DimTab tx = TX[i], ty = TY[j];
    const float ax = tx.a, ay = ty.a;
    const int b0 = ty.cI * C.nx + tx.cI;
    float v = ax * ay * dc[b0];
    if (ax < 1.0f)
      v += (1.0f - ax) * ay * dc[b0 + 1];
    if (ay < 1.0f) {
      v += ax * (1.0f - ay) * dc[b0 + C.nx];
      if (ax < 1.0f)
        v += (1.0f - ax) * (1.0f - ay) * dc[b0 + C.nx + 1];
    }
    d[k] += v;
  }
  team.sync();
}

// the levels from lo down to the coarsest and back (V-cycle below level lo)
template <typename R, typename T>
__device__ __forceinline__ void mg_lower(const DevFieldArgs<R> &a, const MGShared &m, int lo,
                                         T team) {
  const int nl = a.n_levels;
  for (int l = lo; l < nl - 1; ++l) {
    for (int s = 0; s < a.nu_pre; ++s) {
      mg_sweep_c(a, m, l, 0, team);
      mg_sweep_c(a, m, l, 1, team);
    }
    mg_residual_c(a, m, l, m.res, team);
  }
  for (int s = 0; s < a.nu_coarse; ++s) {
    mg_sweep_c(a, m, nl - 1, 0, team);
    mg_sweep_c(a, m, nl - 1, 1, team);
  }
  for (int l = nl - 2; l >= lo; --l) {
    mg_prolong(a, m, l, team);
    for (int s = 0; s < a.nu_post; ++s) {
      mg_sweep_c(a, m, l, 1, team);
      mg_sweep_c(a, m, l, 0, team);
    }
  }
}

__device__ __forceinline__ void nb0(int k, int nx, int ny, int &il, int &ir, int &jd, int &ju) {
  const int j = idiv(k, nx), i = k - j * nx;
  il = i > 0 ? k - 1 : k + 1;
  ir = i < nx - 1 ? k + 1 : k - 1;
  jd = j > 0 ? k - nx : k + nx;
  ju = j < ny - 1 ? k + nx : k - nx;
}

// Fine-grid passes: thread t owns column (t mod 2^cw) of a block of
// 1024 / 2^cw rows; pass p covers the block's rows shifted by p blocks.  The
// index maths is shifts and masks.  A slot off the grid reads and writes
// only the thread's own LDS scratch word (k and the neighbours all point
// there; kg, its global index, is node 0), so no branch separates the loads
// of a pass from its stores.
struct FineSlot {
  int kg, k, il, ir, jd, ju;
  bool ok;
};
__device__ __forceinline__ FineSlot fine_slot(int i, int j, int nx, int ny, int dummy) {
  FineSlot f;
  f.ok = (i < nx) & (j < ny);
  // bitwise selects: the compiler turns nested conditionals into branches
  const int on = -(int)f.ok, off = ~on;
  const int k = j * nx + i;
  f.kg = k & on;
  f.k = (k & on) | (dummy & off);
  f.il = ((k - 1 + 2 * (int)(i == 0)) & on) | (dummy & off);
  f.ir = ((k + 1 - 2 * (int)(i == nx - 1)) & on) | (dummy & off);
  f.jd = ((k - nx + 2 * nx * (int)(j == 0)) & on) | (dummy & off);
  f.ju = ((k + nx - 2 * nx * (int)(j == ny - 1)) & on) | (dummy & off);
  return f;
}
// slot of pass p in the half-row map of colour c (node pairs)
template <int c>
__device__ __forceinline__ FineSlot half_slot(int p, int cw_log2, int nx, int ny, int dummy) {
  const int j = ((int)threadIdx.x >> cw_log2) + p * (kDevThreads >> cw_log2);
  const int i = 2 * ((int)threadIdx.x & ((1 << cw_log2) - 1)) + ((j + c) & 1);
  return fine_slot(i, j, nx, ny, dummy);
}
// slot of pass p in the full-row map
__device__ __forceinline__ FineSlot full_slot(int p, int cw_log2, int nx, int ny, int dummy) {
  const int j = ((int)threadIdx.x >> cw_log2) + p * (kDevThreads >> cw_log2);
  const int i = (int)threadIdx.x & ((1 << cw_log2) - 1);
  return fine_slot(i, j, nx, ny, dummy);
}

// one colour of red-black Gauss-Seidel on the uniform fine grid; the
// right-hand side and diagonal shift (interleaved in global memory,
// Dirichlet nodes carrying (0, kBigShift)) of four passes are loaded before
// their updates
template <int c, typename R>
__device__ __forceinline__ void mg_sweep0(const DevFieldArgs<R> &a, const MGShared &m) {
  float *d = m.d;
  const float hfx = m.hfx, hfy = m.hfy, diag = 2.0f * (m.hfx + m.hfy);
  const int nx = a.nx, ny = a.ny, cw = a.hcw_log2, dummy = a.lds_dummy + (int)threadIdx.x;
  const EMC_GLOBAL float2 *rs0 = glb(a.rs0);
  for (int p0 = 0; p0 < a.nph; p0 += 4) {
    FineSlot f[4];
    float2 v[4];
    float nv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      f[u] = half_slot<c>(p0 + u, cw, nx, ny, dummy);
      v[u] = rs0[f[u].kg];
    }
    // all reads before the stores (the compiler cannot see that the stores
    // never alias the other colour's nodes)
#pragma unroll
    for (int u = 0; u < 4; ++u)
      nv[u] = (v[u].x + hfx * (d[f[u].il] + d[f[u].ir]) + hfy * (d[f[u].jd] + d[f[u].ju])) *
              __builtin_amdgcn_rcpf(diag + v[u].y);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (f[u].ok)
        d[f[u].k] = nv[u];
  }
  __syncthreads();
}

// d_0 += P d_1 on the fine grid in the full-row map: the thread's column
// fixes its x interpolation; Dirichlet nodes are skipped
template <typename R>
__device__ __forceinline__ void mg_prolong0(const DevFieldArgs<R> &a, const MGShared &m) {
  const MGLevel C = a.lev[1];
  float *d = m.d;
  const float *dc = m.d + C.off;
  const DimTab *TX = m.tab + a.lev[0].tx, *TY = m.tab + a.lev[0].ty;
  const int nx = a.nx, ny = a.ny, cw = a.fcw_log2, dummy = a.lds_dummy + (int)threadIdx.x;
  const int i = (int)threadIdx.x & ((1 << cw) - 1);
  const DimTab tx = TX[i < nx ? i : 0];
  const float ax = tx.a;
  const int dx = ax < 1.0f ? 1 : 0;
  for (int p = 0; p < a.npf; ++p) {
    const FineSlot f = full_slot(p, cw, nx, ny, dummy);
    const int j = ((int)threadIdx.x >> cw) + p * (kDevThreads >> cw);
    const DimTab ty = TY[j < ny ? j : 0];
    const float ay = ty.a;
    const int b0 = ty.cI * C.nx + tx.cI;
    const int dy = ay < 1.0f ? C.nx : 0;
    const float v = ax * ay * dc[b0] + (1.0f - ax) * ay * dc[b0 + dx] +
                    ax * (1.0f - ay) * dc[b0 + dy] + (1.0f - ax) * (1.0f - ay) * dc[b0 + dx + dy];
    d[f.k] += glb(a.dmask)[f.kg] ? 0.0f : v;
  }
  __syncthreads();
}

// exact solve on the coarse level lc: d = A_lc^-1 r with the inverse built
// on the host (zero rows and columns at Dirichlet nodes, row-major); four
// threads per row, each summing a contiguous quarter
template <typename R>
__device__ __forceinline__ void mg_direct(const DevFieldArgs<R> &a, const MGShared &m, int lc) {
  const MGLevel C = a.lev[lc];
  const int n = C.nx * C.ny;
  const float *r = m.rc + (C.off - m.n0);
  float *d = m.d + C.off;
  const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
  const int seg = (n + 3) >> 2;
  float s = 0.0f;
  if (row < n) {
    const EMC_GLOBAL float *A = glb(a.cinv) + (long long)row * n;
    const int j1 = (part + 1) * seg < n ? (part + 1) * seg : n;
#pragma unroll 8
    for (int j = part * seg; j < j1; ++j)
      s += A[j] * r[j];
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if (row < n && part == 0)
    d[row] = s;
  __syncthreads();
}

// z = V-cycle(r) into m.d[0 .. n0)
template <typename R>
__device__ __forceinline__ void mg_vcycle(const DevFieldArgs<R> &a, const MGShared &m) {
  const int nl = a.n_levels;
  for (int k = threadIdx.x; k < a.nn; k += kDevThreads)
    m.d[k] = 0.0f;
  __syncthreads();
  if (nl == 1) {
    for (int s = 0; s < a.nu_coarse; ++s) {
      mg_sweep0<0>(a, m);
      mg_sweep0<1>(a, m);
    }
    return;
  }
  for (int s = 0; s < a.nu_pre; ++s) {
    mg_sweep0<0>(a, m);
    mg_sweep0<1>(a, m);
  }
  {  // level-0 residual, restricted to level 1
    const float hfx = m.hfx, hfy = m.hfy, diag = 2.0f * (m.hfx + m.hfy);
    const int nx = a.nx, ny = a.ny, cw = a.fcw_log2, dummy = a.lds_dummy + (int)threadIdx.x;
    const EMC_GLOBAL float2 *rs0 = glb(a.rs0);
    for (int p0 = 0; p0 < a.npf; p0 += 4) {
      FineSlot f[4];
      float2 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        f[u] = full_slot(p0 + u, cw, nx, ny, dummy);
        v[u] = rs0[f[u].kg];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float rk =
            v[u].x - ((diag + v[u].y) * m.d[f[u].k] -
                      (hfx * (m.d[f[u].il] + m.d[f[u].ir]) + hfy * (m.d[f[u].jd] + m.d[f[u].ju])));
        // m.res[k] (m.res = m.d + lds_res), the scratch word off the grid
        m.d[f[u].ok ? a.lds_res + f[u].k : f[u].k] = v[u].y >= kBigShift ? 0.0f : rk;
      }
    }
    __syncthreads();
    mg_restrict_to(a, m, 0, m.res, m.rc + (a.lev[1].off - m.n0), true, BlockTeam());
  }
  // the coarse level below which the hierarchy is replaced by its exact
  // inverse (when the Jacobian shift is negligible there), or handed to one
  // wave
  const int lc = a.direct_level;
  const bool direct = lc > 0 && lc < nl && !a.equilibrium;
  const int lw = direct ? lc : (a.wave_level < nl ? a.wave_level : nl);
  for (int l = 1; l < lw && l < nl - 1; ++l) {
    for (int s = 0; s < a.nu_pre; ++s) {
      mg_sweep_c(a, m, l, 0, BlockTeam());
      mg_sweep_c(a, m, l, 1, BlockTeam());
    }
    mg_residual_c(a, m, l, m.res, BlockTeam());
  }
  if (direct) {
    mg_direct(a, m, lc);
  } else if (lw < nl) {
    if (threadIdx.x < 64)
      mg_lower(a, m, lw, WaveTeam());
    __syncthreads();
  } else {
    for (int s = 0; s < a.nu_coarse; ++s) {
      mg_sweep_c(a, m, nl - 1, 0, BlockTeam());
      mg_sweep_c(a, m, nl - 1, 1, BlockTeam());
    }
  }
  for (int l = (lw < nl ? lw : nl - 1) - 1; l >= 1; --l) {
    mg_prolong(a, m, l, BlockTeam());
    for (int s = 0; s < a.nu_post; ++s) {
      mg_sweep_c(a, m, l, 1, BlockTeam());
      mg_sweep_c(a, m, l, 0, BlockTeam());
    }
  }
  mg_prolong0(a, m);
  for (int s = 0; s < a.nu_post; ++s) {
    mg_sweep0<1>(a, m);
    mg_sweep0<0>(a, m);
  }
}

// diagonal shifts of the coarse levels from the level-0 one (lumped
// Galerkin P^T diag(sigma) P)
template <typename R>
__device__ __forceinline__ void mg_shift_down(const DevFieldArgs<R> &a, const MGShared &m) {
  const int nl = a.n_levels;
  if (nl == 1)
    return;
  for (int k = threadIdx.x; k < a.nn; k += kDevThreads) {
    const float sk = a.rs0[k].y;
    m.res[k] = sk >= kBigShift ? 0.0f : sk;
  }
  __syncthreads();
  mg_restrict_to(a, m, 0, m.res, m.sc + (a.lev[1].off - m.n0), false, BlockTeam(), true);
  for (int l = 1; l < nl - 1; ++l) {
    // the restriction input: the finer level's shift without its Dirichlet
    // values
    const MGLevel L = a.lev[l];
    const float *sf = m.sc + (L.off - m.n0);
    for (int k = threadIdx.x; k < L.nx * L.ny; k += kDevThreads)
      m.res[k] = sf[k] >= kBigShift ? 0.0f : sf[k];
    __syncthreads();
    mg_restrict_to(a, m, l, m.res, m.sc + (a.lev[l + 1].off - m.n0), false, BlockTeam(), true);
  }
}

// Newton's method on the discrete equations of emcSORSolver
// (PoissonSolver/emcSORSolver.hpp:49-197):
//   F_k = sum_d hF_d (phi_prev + phi_next - 2 phi_k) + hP (p - n + C)
//         + sum_gates g hF_d h_d (V_g - phi_k) = 0
// at every non-ohmic node (mirror neighbours at the boundary), with
// p = exp(-phi) and n = conc (non-equilibrium) or exp(phi) (equilibrium);
// ohmic nodes hold V/V_T + asinh(C/2).  Residuals, the conjugate-gradient
// vectors and the update are FP64; the CG preconditioner is one FP32
// multigrid V-cycle.  Then calcEField (PMSchemes/emcEFieldCalculation.hpp:
// 13-81).
template <typename R, bool WS>
__global__ __launch_bounds__(kDevThreads) void k_dev_field(const DevFieldArgs<R> *__restrict__ args) {
  const DevFieldArgs<R> &a = *args;
  extern __shared__ __align__(16) float lds[];
  // the multigrid data: in LDS, or with WS (grids whose levels exceed the
  // LDS) in a global workspace of the same layout, kept in L2
  float *const mg = WS ? a.mg_ws : lds;
  __shared__ double red[kDevThreads / 64];
  const int tid = threadIdx.x;
  const int nx = a.nx, ny = a.ny, nn = a.nn;
  double *pot = a.pot;
  long long row = 0;
  if (a.record)
    row = a.state->step - a.state->hist_base;
  if (a.record && a.nrpart) {
    double s = 0.0;
    for (int k = tid; k < nn; k += kDevThreads)
      s += a.nrpart[k];
    s = block_sum(s, red);
    if (tid == 0) {
      const double na = (double)a.state->n_alive;
      a.hist.checksum[row] = fabs(s - na) / (na > 1.0 ? na : 1.0);
    }
  }
  // setPoissonInterval (emcSimulation.hpp:84-86): between solves the field
  // of the last solve is kept
  if (a.record && a.interval > 1 && a.state->step % a.interval != 0) {
    if (tid == 0) {
      a.hist.iters[row] = 0;
      a.hist.resid[row] = 0.0;
    }
    return;
  }
  // the Newton corrections before this launch: their checks say how many
  // ran and whether one already meets the stopping rule (the last check
  // follows the last correction; a separate launch over all CUs forms it,
  // which is cheaper than one workgroup here)
  int green_done = 0;
  bool green_ok = false;
  double green_fm = 0.0, green_scale = 1.0;
  if (a.do_poisson && a.green_iters > 0 && a.green_acc &&
      *a.green_flag != (unsigned long long)a.state->step + 1ull) {
    for (int c = 0; c <= a.green_iters && !green_ok; ++c) {
      const double fm_c = __longlong_as_double((long long)a.green_acc[2 * c]);
      double sc_c = __longlong_as_double((long long)a.green_acc[2 * c + 1]);
      sc_c = sc_c > 0.0 ? sc_c : 1.0;
      green_fm = fm_c;
      green_scale = sc_c;
      if (fm_c <= a.tol * sc_c)
        green_ok = true;  // the corrections after this check were skipped
      else if (c < a.green_iters)
        ++green_done;
    }
  }
  if (a.green_iters > 0 && a.green_acc) {
    __syncthreads();  // every lane has read the checks
    if (tid < 2 * (a.green_iters + 1))
      a.green_acc[tid] = 0ull;
  }
  if (a.do_poisson && green_ok && tid == 0) {
    if (a.iters_out)
      *a.iters_out = green_done;
    if (a.resid_out)
      *a.resid_out = green_fm / green_scale;
    if (a.record) {
      a.hist.iters[row] = green_done;
      a.hist.resid[row] = green_fm / green_scale;
    }
  }
  if (a.do_poisson && !green_ok) {
    MGShared m;
    m.d = mg;
    m.res = mg + a.lds_res;
    m.rc = mg + a.lds_rc;
    m.sc = mg + a.lds_sc;
    DimTab *tab = (DimTab *)((unsigned char *)mg + a.lds_tab);
    uint8_t *cmask = (uint8_t *)mg + a.lds_cmask;
    m.tab = tab;
    m.cmask = cmask;
    m.n0 = nn;
    m.hfx = (float)a.hfx;
    m.hfy = (float)a.hfy;
    for (int e = tid; e < a.n_tab; e += kDevThreads)
      tab[e] = a.tab[e];
    for (int e = tid; e < a.n_coarse; e += kDevThreads) {
      const uint8_t mk = a.lmask[nn + e];
      cmask[e] = mk;
      m.sc[e] = mk ? kBigShift : 0.0f;
    }
    if (a.reset_bc)
      for (int k = tid; k < nn; k += kDevThreads)
        if (a.dmask[k])
          pot[k] = a.equilibrium ? a.bc_eq[k] : a.bc_neq[k];
    __syncthreads();
    const double hfx = a.hfx, hfy = a.hfy, hp = a.hp, hfs2 = 2.0 * (a.hfx + a.hfy);
    EMC_GLOBAL double *r64 = glb(a.r64), *x64 = glb(a.x64), *p64 = glb(a.p64),
                      *q64 = glb(a.q64), *s64 = glb(a.s64);
    EMC_GLOBAL double *potg = glb(pot);
    EMC_GLOBAL float2 *rs0 = glb(a.rs0);
    const EMC_GLOBAL double *conc = glb(a.conc), *dopn = glb(a.dopn);
    const EMC_GLOBAL double *gate = a.gate ? glb(a.gate) : nullptr;
    const EMC_GLOBAL float *w = glb(a.w0);  // 1 at active nodes, 0 at Dirichlet and padding
    // control-volume weights (1/2 per boundary dimension): the mirror
    // boundary operator is symmetric in this inner product
    const EMC_GLOBAL float *hw = glb(a.hw0);
    const int np = a.n_pad;  // nn rounded up to kPad nodes
    double scale = 1.0, fm = 0.0;
    int it = 0, vcycles = 0;
    for (;; ++it) {
      // Newton residual F and the Jacobian's diagonal shift
      double lmax = 0.0, lsc = 0.0;
      for (int k0 = tid; k0 < np; k0 += kPad) {
#pragma unroll
        for (int u = 0; u < kPad / kDevThreads; ++u) {
          const int k = k0 + u * kDevThreads;
          const float wk = w[k];
          const int kc = k < nn ? k : nn - 1;
          int il, ir, jd, ju;
          nb0(kc, nx, ny, il, ir, jd, ju);
          const double phi = potg[kc];
          const double p = exp(-phi);
          const double n = a.equilibrium ? exp(phi) : conc[kc];
          const double C = dopn[kc];
          double gt = 0.0, sg = hp * (a.equilibrium ? p + n : p);
          if (gate) {  // gate (Robin) term and its Jacobian shift
            const double gg = gate[3 * kc];
            gt = gate[3 * kc + (a.equilibrium ? 1 : 2)] - gg * phi;
            sg += gg;
          }
          const double F = wk * (hfx * (potg[il] + potg[ir]) + hfy * (potg[jd] + potg[ju]) -
                                 hfs2 * phi + hp * (p - n + C) + gt);
          r64[k] = F;
          s64[k] = sg;
          rs0[k] = make_float2((float)F, wk != 0.0f ? (float)sg : kBigShift);
          const double af = fabs(F);
          lmax = af > lmax ? af : lmax;
          const double sc = wk * hp * (fabs(C) + n + p);
          lsc = sc > lsc ? sc : lsc;
        }
      }
      fm = block_max(lmax, red);
      if (it == 0) {
        scale = block_max(lsc, red);
        scale = scale > 0.0 ? scale : 1.0;
      }
      if (fm <= a.tol * scale || it >= a.max_iter || vcycles >= a.max_vcycles)
        break;
      if (a.equilibrium || it == 0)
        mg_shift_down(a, m);
      // conjugate gradients on A x = F (A = -Jacobian, symmetric in the
      // control-volume inner product), x0 = 0
      // inexact Newton away from the root: the FP32-preconditioned CG
      // solve stops at a tenth of the residual and the system is
      // linearised again (solving each Newton system to the stopping rule
      // takes several times the V-cycles once the contacts' voltages or a
      // p-type region's hole term pull the potential far from the start)
      double target = 0.5 * a.tol * scale;
      if (0.1 * fm > target)
        target = 0.1 * fm;
      double rz_old = 1.0;
      for (int k = tid; k < np; k += kDevThreads) {
        x64[k] = 0.0;
        p64[k] = 0.0;
      }
      for (int pit = 0; pit < a.max_pcg; ++pit) {
        mg_vcycle(a, m);  // z = M^-1 r in m.d[0 .. nn)
        ++vcycles;
        double rz = 0.0;
        for (int k0 = tid; k0 < np; k0 += kPad) {
#pragma unroll
          for (int u = 0; u < kPad / kDevThreads; ++u) {
            const int k = k0 + u * kDevThreads;
            rz += hw[k] * r64[k] * (double)(w[k] * m.d[k < nn ? k : nn - 1]);
          }
        }
        rz = block_sum(rz, red);
        const double beta = pit == 0 ? 0.0 : rz / rz_old;
        rz_old = rz;
        for (int k0 = tid; k0 < np; k0 += kPad) {
#pragma unroll
          for (int u = 0; u < kPad / kDevThreads; ++u) {
            const int k = k0 + u * kDevThreads;
            p64[k] = (double)(w[k] * m.d[k < nn ? k : nn - 1]) + beta * p64[k];
          }
        }
        __syncthreads();
        double pq = 0.0;
        for (int k0 = tid; k0 < np; k0 += kPad) {
#pragma unroll
          for (int u = 0; u < kPad / kDevThreads; ++u) {
            const int k = k0 + u * kDevThreads;
            const int kc = k < nn ? k : nn - 1;
            int il, ir, jd, ju;
            nb0(kc, nx, ny, il, ir, jd, ju);
            const double pk = p64[k];
            const double qq = w[k] * ((hfs2 + s64[k]) * pk -
                                      (hfx * (p64[il] + p64[ir]) + hfy * (p64[jd] + p64[ju])));
            q64[k] = qq;
            pq += hw[k] * pk * qq;
          }
        }
        pq = block_sum(pq, red);
        const double alpha = pq != 0.0 ? rz / pq : 0.0;
        double rmax = 0.0;
        for (int k0 = tid; k0 < np; k0 += kPad) {
#pragma unroll
          for (int u = 0; u < kPad / kDevThreads; ++u) {
            const int k = k0 + u * kDevThreads;
            x64[k] += alpha * p64[k];
            const double rn = r64[k] - alpha * q64[k];
            r64[k] = rn;
            rs0[k].x = (float)rn;
            const double ar = fabs(rn);
            rmax = ar > rmax ? ar : rmax;
          }
        }
        rmax = block_max(rmax, red);
        if (rmax <= target || vcycles >= a.max_vcycles)
          break;
      }
      for (int k = tid; k < nn; k += kDevThreads) {
        // damped Newton: a correction beyond 2 V_T is clipped (from the
        // charge-neutral guess, and where new contact voltages move the
        // potential across a junction, the hole term exp(-phi) would
        // otherwise overshoot); near the root the corrections are far
        // smaller and the step is Newton's
        double d = x64[k];  // zero at Dirichlet nodes
        d = d > 2.0 ? 2.0 : d;
        d = d < -2.0 ? -2.0 : d;
        potg[k] += d;
      }
      __syncthreads();
    }
    if (tid == 0) {
      if (a.iters_out)
        *a.iters_out = vcycles;
      if (a.resid_out)
        *a.resid_out = fm / scale;
      if (a.record) {
        a.hist.iters[row] = vcycles + green_done;
        a.hist.resid[row] = fm / scale;
      }
    }
    __syncthreads();
  }
  if (a.do_efield && !(a.efield_if_solved && green_ok)) {
#pragma clang fp contract(off)
    const bool nec = a.pm == EMCGPU_PM_NEC;
    for (int k = tid; k < nn; k += kDevThreads) {
      const int j = idiv(k, nx), i = k - j * nx;
      if (i != 0 && i != nx - 1)
        a.ex[k] = nec ? (pot[k] - pot[k + 1]) * a.vt / a.hxm
                      : (pot[k - 1] - pot[k + 1]) * a.vt / (2 * a.hxm);
      if (j != 0 && j != ny - 1)
        a.ey[k] = nec ? (pot[k] - pot[k + nx]) * a.vt / a.hym
                      : (pot[k - nx] - pot[k + nx]) * a.vt / (2 * a.hym);
    }
    __syncthreads();
    // setEFieldBoundaryValues: a contact copies the inner neighbour, any
    // other boundary node is zero
    for (int k = tid; k < nn; k += kDevThreads) {
      const int j = idiv(k, nx), i = k - j * nx;
      const int bx = a.bcx[k], by = a.bcy[k];
      if (bx != -2)
        a.ex[k] = bx == -1 ? 0.0 : a.ex[i == 0 ? k + 1 : k - 1];
      if (by != -2)
        a.ey[k] = by == -1 ? 0.0 : a.ey[j == 0 ? k + nx : k - nx];
    }
    __syncthreads();
    if (a.acc)
      for (int k = tid; k < nn; k += kDevThreads) {
        a.acc[2 * k] = (R)(a.ex[k] * a.acc_scale);
        a.acc[2 * k + 1] = (R)(a.ey[k] * a.acc_scale);
      }
  }
}

// The E-field and force grid of k_dev_field over all CUs, one node per
// lane, for grids of at least 3 x 3 nodes (a contact node copies its inner
// neighbour, which is then an inner node whose value is formed here
// directly; the formulas and their rounding are those of k_dev_field).
template <typename R>
__global__ __launch_bounds__(256) void k_dev_efield(const DevFieldArgs<R> *__restrict__ args) {
#pragma clang fp contract(off)
  const DevFieldArgs<R> &a = *args;
  if (a.record && a.interval > 1 && a.state->step % a.interval != 0)
    return;  // the field of the last solve is kept
  const int nx = a.nx, nn = a.nn;
  const int k = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (k >= nn)
    return;
  const double *pot = a.pot;
  const bool nec = a.pm == EMCGPU_PM_NEC;
  const int j = idiv(k, nx), i = k - j * nx;
  auto fx = [&](int q) {
    return nec ? (pot[q] - pot[q + 1]) * a.vt / a.hxm
               : (pot[q - 1] - pot[q + 1]) * a.vt / (2 * a.hxm);
  };
  auto fy = [&](int q) {
    return nec ? (pot[q] - pot[q + nx]) * a.vt / a.hym
               : (pot[q - nx] - pot[q + nx]) * a.vt / (2 * a.hym);
  };
  const int bx = a.bcx[k], by = a.bcy[k];
  double ex = a.ex[k], ey = a.ey[k];
  if (bx == -2)
    ex = fx(k);
  else
    ex = bx == -1 ? 0.0 : fx(i == 0 ? k + 1 : k - 1);
  if (by == -2)
    ey = fy(k);
  else
    ey = by == -1 ? 0.0 : fy(j == 0 ? k + nx : k - nx);
  a.ex[k] = ex;
  a.ey[k] = ey;
  if (a.acc) {
    a.acc[2 * k] = (R)(ex * a.acc_scale);
    a.acc[2 * k + 1] = (R)(ey * a.acc_scale);
  }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
template <typename R>
hipError_t dev_launch_efield(const DevFieldArgs<R> *args_dev, int nn, hipStream_t st) {
  hipLaunchKernelGGL((k_dev_efield<R>), dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st,
                     args_dev);
  return hipGetLastError();
}

// the arguments live in device memory (args_dev, written beforehand by
// the caller): dynamic indexing of the level table would otherwise copy the
// whole block to scratch
template <typename R>
hipError_t dev_launch_field(const DevFieldArgs<R> *args_dev, size_t lds, bool ws, hipStream_t st) {
  if (ws) {  // the multigrid data in the global workspace (args_dev->mg_ws)
    hipLaunchKernelGGL((k_dev_field<R, true>), dim3(1), dim3(kDevThreads), 0, st, args_dev);
    return hipGetLastError();
  }
  if (lds > (size_t)kDevMaxLds)
    return hipErrorInvalidValue;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_dev_field<R, false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e;
    }
  }
  hipLaunchKernelGGL((k_dev_field<R, false>), dim3(1), dim3(kDevThreads), lds, st, args_dev);
  return hipGetLastError();
}

template <typename R>
hipError_t dev_launch_step(const DevStepArgs<R> &a, const DevContactArgs<R> &ca, int blocks,
                           int n_seg, size_t lds, hipStream_t st) {
  if (lds > (size_t)kDevMaxLds)
    return hipErrorInvalidValue;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_dev_step<R>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e;
    }
  }
  hipLaunchKernelGGL((k_dev_step<R>), dim3(blocks), dim3(DevStepShape<R>::threads), lds, st, a);
  const int n_rank = n_seg / kDevStepWaves;
  const int node_blocks = (ca.g.nn + kDevStepThreads - 1) / kDevStepThreads;
  hipLaunchKernelGGL((k_dev_rank_rows<R>), dim3((unsigned)(n_rank + node_blocks * ca.n_chunks)),
                     dim3(kDevStepThreads), sizeof(uint32_t) * kDevStepWaves * (size_t)a.n_slots,
                     st, a, ca, n_rank);
  return hipGetLastError();
}

template <typename R>
hipError_t dev_launch_drift(const DevStepArgs<R> &a, int mode, hipStream_t st) {
  const dim3 g((unsigned)((a.n + 255) / 256)), b(256);
  if (mode == 1)
    hipLaunchKernelGGL((k_dev_drift<R, 1>), g, b, 0, st, a);
  else
    hipLaunchKernelGGL((k_dev_drift<R, 0>), g, b, 0, st, a);
  return hipGetLastError();
}

template <typename R>
hipError_t dev_launch_contacts(const DevContactArgs<R> &a, long long max_inj, hipStream_t st) {
  hipLaunchKernelGGL((k_dev_contact_scan<R>), dim3(a.n_slots), dim3(256), 0, st, a);
  const long long nA = (a.n_segments + 3) / 4, nB = (max_inj + 255) / 256;
  hipLaunchKernelGGL((k_dev_resolve<R>), dim3((unsigned)(nA + nB)), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <typename R>
hipError_t dev_launch_reduce(const DevContactArgs<R> &a, int commit, hipStream_t st) {
  hipLaunchKernelGGL((k_dev_reduce<R>), dim3((a.g.nn + 255) / 256), dim3(256), 0, st, a, commit);
  return hipGetLastError();
}

template <typename R>
hipError_t dev_launch_init(const DevContactArgs<R> &a, const long long *prefix, long long n,
                           hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL((k_dev_init<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a,
                     prefix, n);
  return hipGetLastError();
}

template <typename R>
hipError_t dev_launch_compact(const DevParticles<R> &src, const DevParticles<R> &dst, int *blk,
                              long long n_total, DevState *state, hipStream_t st) {
  const long long nb = (n_total + kCompactTile - 1) / kCompactTile;
  if (nb <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_dev_count_alive, dim3((unsigned)nb), dim3(256), 0, st, src.vs, state, blk);
  hipLaunchKernelGGL(k_dev_scan_blocks, dim3(1), dim3(1024), 0, st, blk, (int)nb, state);
  hipLaunchKernelGGL((k_dev_compact<R>), dim3((unsigned)nb), dim3(256), 0, st, src, dst, blk,
                     state);
  hipLaunchKernelGGL(k_dev_compact_commit, dim3(1), dim3(1), 0, st, state);
  return hipGetLastError();
}

hipError_t dev_launch_mark(DevState *state, hipStream_t st) {
  hipLaunchKernelGGL(k_dev_mark, dim3(1), dim3(1), 0, st, state);
  return hipGetLastError();
}

// resident step blocks per CU at the given dynamic LDS
template <typename R>
int dev_step_blocks_per_cu(size_t lds) {
  int n = 0;
  if (lds > 65536 &&
      hipFuncSetAttribute((const void *)k_dev_step<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_dev_step<R>, DevStepShape<R>::threads,
                                                   lds) !=
      hipSuccess)
    return 1;
  return n;
}

hipError_t dev_launch_zero(double *p, long long n, hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_dev_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}
template <typename R>
hipError_t dev_launch_zero_r(R *p, long long n, hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL((k_dev_zero_r<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}
hipError_t dev_launch_zero_i(int *p, long long n, hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_dev_zero_i, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t dev_launch_assign_pos(const DevGeom<double> &g, const double *pos, long long n,
                                 double *grid, hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_dev_assign_pos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, pos,
                     n, grid);
  return hipGetLastError();
}
hipError_t dev_launch_interp_pos(const DevGeom<double> &g, const double *pos, long long n,
                                 double charge, double *force, hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_dev_interp_pos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, pos,
                     n, charge, force);
  return hipGetLastError();
}

hipError_t dev_launch_pm3_assign(const PMGrid3 &g, const double *pos, long long n, double *grid,
                                 hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_pm3_assign, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, pos, n,
                     grid);
  return hipGetLastError();
}
hipError_t dev_launch_pm3_interp(const PMGrid3 &g, const double *ex, const double *ey,
                                 const double *ez, const double *pos, long long n, double charge,
                                 double *force, hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_pm3_interp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, ex, ey,
                     ez, pos, n, charge, force);
  return hipGetLastError();
}
hipError_t dev_launch_pm3_efield(const PMGrid3 &g, double vt, const signed char *bc,
                                 const double *pot, double *ex, double *ey, double *ez,
                                 hipStream_t st) {
  const long long nn = (long long)g.nx * g.ny * g.nz;
  const dim3 grid((unsigned)((nn + 255) / 256));
  hipLaunchKernelGGL(k_pm3_efield, grid, dim3(256), 0, st, g, vt, pot, ex, ey, ez);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_pm3_faces, grid, dim3(256), 0, st, g, bc, ex, ey, ez);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// dense Newton corrections of the per-step Poisson solve
// ---------------------------------------------------------------------------
// Newton residual F of the non-equilibrium equation at every node (the
// stencil of k_dev_field), g = H F, and residual check c: max |F| and the
// stopping rule's scale into acc[c].  A node whose hole term h^2 p is not
// negligible against the diagonal marks the step, so its corrections are
// skipped and the multigrid solve works alone.
__device__ __forceinline__ bool green_converged(const GreenArgs &a, int c) {
  const double fm = __longlong_as_double((long long)a.acc[2 * c]);
  const double sc = __longlong_as_double((long long)a.acc[2 * c + 1]);
  return fm <= a.tol * (sc > 0.0 ? sc : 1.0);
}

__global__ __launch_bounds__(256) void k_green_res(GreenArgs a, int c) {
#pragma clang fp contract(off)
  __shared__ double red[2][4];
  if (a.interval > 1 && a.state->step % a.interval != 0)
    return;
  const int k = (int)blockIdx.x * 256 + (int)threadIdx.x;
  double af = 0.0, sc = 0.0;
  if (k < a.nn) {
    const double wk = a.w0[k];
    int il, ir, jd, ju;
    nb0(k, a.nx, a.ny, il, ir, jd, ju);
    const double *pot = a.pot;
    const double phi = pot[k];
    const double p = exp(-phi);
    const double n = a.conc[k], C = a.dopn[k];
    const double gt = a.gate ? a.gate[3 * k + 2] - a.gate[3 * k] * phi : 0.0;
    const double F = wk * (a.hfx * (pot[il] + pot[ir]) + a.hfy * (pot[jd] + pot[ju]) -
                           2.0 * (a.hfx + a.hfy) * phi + a.hp * (p - n + C) + gt);
    a.g[k] = a.use_fd ? F : (double)a.hw0[k] * F;
    af = fabs(F);
    sc = wk * a.hp * (fabs(C) + n + p);
    if (wk != 0.0 && a.hp * p > a.shift_max)
      atomicMax(a.flag, (unsigned long long)a.state->step + 1ull);
    if (a.fuse_efield && c == a.iters) {
      // calcEField of the corrected potential (k_dev_efield)
      const int nx = a.nx, j = idiv(k, nx), i = k - j * nx;
      const bool nec = a.pm == EMCGPU_PM_NEC;
      auto fx = [&](int q) {
        return nec ? (pot[q] - pot[q + 1]) * a.vt / a.hxm
                   : (pot[q - 1] - pot[q + 1]) * a.vt / (2 * a.hxm);
      };
      auto fy = [&](int q) {
        return nec ? (pot[q] - pot[q + nx]) * a.vt / a.hym
                   : (pot[q - nx] - pot[q + nx]) * a.vt / (2 * a.hym);
      };
      const int bx = a.bcx[k], by = a.bcy[k];
      double ex, ey;
      if (bx == -2)
        ex = fx(k);
      else
        ex = bx == -1 ? 0.0 : fx(i == 0 ? k + 1 : k - 1);
      if (by == -2)
        ey = fy(k);
      else
        ey = by == -1 ? 0.0 : fy(j == 0 ? k + nx : k - nx);
      a.ex[k] = ex;
      a.ey[k] = ey;
      if (a.acc_fp64) {
        ((double *)a.force)[2 * k] = ex * a.acc_scale;
        ((double *)a.force)[2 * k + 1] = ey * a.acc_scale;
      } else {
        ((float *)a.force)[2 * k] = (float)(ex * a.acc_scale);
        ((float *)a.force)[2 * k + 1] = (float)(ey * a.acc_scale);
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double x = __shfl_xor(af, o, 64), y = __shfl_xor(sc, o, 64);
    af = x > af ? x : af;
    sc = y > sc ? y : sc;
  }
  const int wave = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wave] = af;
    red[1][wave] = sc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      af = red[0][w] > af ? red[0][w] : af;
      sc = red[1][w] > sc ? red[1][w] : sc;
    }
    // non-negative doubles order like their bit patterns
    atomicMax(a.acc + 2 * c, (unsigned long long)__double_as_longlong(af));
    atomicMax(a.acc + 2 * c + 1, (unsigned long long)__double_as_longlong(sc));
  }
}

// x = G g with FP64 sums, pot += x at the active nodes.  g is staged in LDS
// (FP32, the precision of G); each wave takes one row of G at a time and
// reads it as coalesced float4 (the kernel streams G from HBM once).
typedef float green_f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kGreenThreads) void k_green_apply(GreenArgs a, int c) {
  extern __shared__ __align__(16) float gs[];  // [ld]
  if (a.interval > 1 && a.state->step % a.interval != 0)
    return;
  // the same values for every block: the whole grid leaves together
  if (*a.flag == (unsigned long long)a.state->step + 1ull || green_converged(a, c))
    return;
  for (int j = threadIdx.x; j < a.ld; j += kGreenThreads)
    gs[j] = j < a.nn ? (float)a.g[j] : 0.0f;
  __syncthreads();
  const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int rows_per_block = kGreenThreads / 64;
  const int q4 = a.ld / 4;  // float4 per row, a multiple of 64
  const green_f4 *g4 = reinterpret_cast<const green_f4 *>(gs);
  for (int row = (int)blockIdx.x * rows_per_block + wave; row < a.nn;
       row += (int)gridDim.x * rows_per_block) {
    if (a.w0[row] == 0.0f)
      continue;  // the same for the whole wave
    const green_f4 *G = reinterpret_cast<const green_f4 *>(a.green + (size_t)row * a.ld);
    // four row segments in flight per lane
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    int q = lane;
    for (; q + 192 < q4; q += 256) {
      green_f4 m[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        m[u] = __builtin_nontemporal_load(G + q + 64 * u);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const green_f4 v = g4[q + 64 * u];
        acc[u] = fma((double)m[u].x, (double)v.x, acc[u]);
        acc[u] = fma((double)m[u].y, (double)v.y, acc[u]);
        acc[u] = fma((double)m[u].z, (double)v.z, acc[u]);
        acc[u] = fma((double)m[u].w, (double)v.w, acc[u]);
      }
    }
    for (; q < q4; q += 64) {
      const green_f4 m0 = G[q];
      const green_f4 v = g4[q];
      acc[0] = fma((double)m0.x, (double)v.x, acc[0]);
      acc[0] = fma((double)m0.y, (double)v.y, acc[0]);
      acc[0] = fma((double)m0.z, (double)v.z, acc[0]);
      acc[0] = fma((double)m0.w, (double)v.w, acc[0]);
    }
    double acc_row = (acc[0] + acc[1]) + (acc[2] + acc[3]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      acc_row += __shfl_xor(acc_row, o, 64);
    if (lane == 0)
      a.pot[row] += acc_row;
  }
}

// dynamic LDS of the apply kernel beyond the default 64 KB
hipError_t dev_green_lds(size_t lds) {
  if (lds <= 64 * 1024)
    return hipSuccess;
  return hipFuncSetAttribute((const void *)k_green_apply,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// Newton correction by the separable inverse.  Both launches run a block per
// (active row j, chunk of kFdChunk columns) and are bound by the latency of
// dependent L2 round trips, so every sum keeps eight loads in flight per
// lane and is split over as many threads as the block has: the row of the
// y transform (all mx columns, in LDS) takes P = kFdThreads / mx threads
// per column, each summing a P-th of the rows; the chunk's x sums take one
// lane per output column and the block's waves each a slice of the rows.
constexpr int kFdChunk = 64;
constexpr int kFdWaves = kFdThreads / 64;

// sum over r in [r0, r1) of m[r * ld] * v(r), eight loads in flight
template <typename V>
__device__ __forceinline__ double fd_dot8(const double *m, size_t ld, int r0, int r1, V v) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  int r = r0;
  for (; r + 8 <= r1; r += 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      x[u] = m[(size_t)(r + u) * ld];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      s[u & 3] = fma(x[u], v(r + u), s[u & 3]);
  }
  for (; r < r1; ++r)
    s[0] = fma(m[(size_t)r * ld], v(r), s[0]);
  return (s[0] + s[1]) + (s[2] + s[3]);
}

// row[q] = sum_b w[b] in(b, q) for the mx columns q, P threads per column
// (part: max(mx, kFdThreads) doubles of LDS)
template <typename IN>
__device__ __forceinline__ void fd_row(const double *w, int my, int mx, double *row, double *part,
                                       IN in) {
  const int P = max(1, min(kFdThreads / mx, 8));
  const int span = (my + P - 1) / P;
  for (int e = threadIdx.x; e < P * mx; e += kFdThreads) {
    const int p = e / mx, q = e - p * mx;
    const int b0 = p * span, b1 = min(my, b0 + span);
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int b = b0;
    for (; b + 8 <= b1; b += 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        x[u] = in(b + u, q);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        s[u & 3] = fma(w[b + u], x[u], s[u & 3]);
    }
    for (; b < b1; ++b)
      s[0] = fma(w[b], in(b, q), s[0]);
    part[e] = (s[0] + s[1]) + (s[2] + s[3]);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < mx; q += kFdThreads) {
    double t = 0.0;
    for (int p = 0; p < P; ++p)
      t += part[p * mx + q];
    row[q] = t;
  }
  __syncthreads();
}

// the chunk's x sums: out = sum_r M[r][k] row[r] for lane k of the chunk,
// each wave a slice of the rows; the result is valid in wave 0
__device__ __forceinline__ double fd_chunk(const double *M, const double *row, double *part,
                                           int mx, int k) {
  const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int span = (mx + kFdWaves - 1) / kFdWaves, r0 = w * span, r1 = min(mx, r0 + span);
  part[w * kFdChunk + lane] =
      k < mx ? fd_dot8(M + k, (size_t)mx, r0, r1, [&](int r) { return row[r]; }) : 0.0;
  __syncthreads();
  double t = 0.0;
  if (w == 0)
    for (int v = 0; v < kFdWaves; ++v)
      t += part[v * kFdChunk + lane];
  return t;
}

// t[j][k] = sum_q Px[k][q] (sum_b Py[j][b] F[b][q]) / (lx[k] + ly[j])
__global__ __launch_bounds__(kFdThreads) void k_fd_fwd(GreenArgs a, int c) {
  extern __shared__ double fd_lds[];  // [mx] row, [max(mx, kFdThreads)] partial sums
  if (a.interval > 1 && a.state->step % a.interval != 0)
    return;
  if (*a.flag == (unsigned long long)a.state->step + 1ull || green_converged(a, c))
    return;
  const FdOps &f = a.fd;
  const int j = blockIdx.y, mx = f.mx, my = f.my, nx = a.nx;
  double *row = fd_lds, *part = fd_lds + mx;
  fd_row(f.py + (size_t)j * my, my, mx, row, part,
         [&](int b, int q) { return a.g[(size_t)f.iy[b] * nx + f.ix[q]]; });
  const int k = (int)blockIdx.x * kFdChunk + ((int)threadIdx.x & 63);
  const double t = fd_chunk(f.pxt, row, part, mx, k);
  if (threadIdx.x < 64 && k < mx)
    f.t[(size_t)j * mx + k] = t * f.inv_den[(size_t)j * mx + k];
}

// pot[j][q] += sum_k Bx[q][k] (sum_l By[j][l] t[l][k])
__global__ __launch_bounds__(kFdThreads) void k_fd_bwd(GreenArgs a, int c) {
  extern __shared__ double fd_lds[];  // [mx] row, [max(mx, kFdThreads)] partial sums
  if (a.interval > 1 && a.state->step % a.interval != 0)
    return;
  if (*a.flag == (unsigned long long)a.state->step + 1ull || green_converged(a, c))
    return;
  const FdOps &f = a.fd;
  const int j = blockIdx.y, mx = f.mx, my = f.my;
  double *row = fd_lds, *part = fd_lds + mx;
  fd_row(f.by + (size_t)j * my, my, mx, row, part,
         [&](int l, int k) { return f.t[(size_t)l * mx + k]; });
  const int q = (int)blockIdx.x * kFdChunk + ((int)threadIdx.x & 63);
  const double x = fd_chunk(f.bxt, row, part, mx, q);
  if (threadIdx.x < 64 && q < mx)
    a.pot[(size_t)f.iy[j] * a.nx + f.ix[q]] += x;
}

// residual check, correction (skipped once a check meets the stopping
// rule), ..., and a last check after the last correction
hipError_t dev_launch_green(const GreenArgs &a, int iters, int blocks, hipStream_t st) {
  const unsigned rb = (unsigned)((a.nn + 255) / 256);
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_green_res, dim3(rb), dim3(256), 0, st, a, it);
    if (a.use_fd) {
      const size_t lds = (size_t)(a.fd.mx + max(a.fd.mx, kFdThreads)) * sizeof(double);
      const dim3 grid((unsigned)((a.fd.mx + kFdChunk - 1) / kFdChunk), (unsigned)a.fd.my);
      hipLaunchKernelGGL(k_fd_fwd, grid, dim3(kFdThreads), lds, st, a, it);
      hipLaunchKernelGGL(k_fd_bwd, grid, dim3(kFdThreads), lds, st, a, it);
    } else {
      hipLaunchKernelGGL(k_green_apply, dim3((unsigned)blocks), dim3(kGreenThreads),
                         (size_t)a.ld * sizeof(float), st, a, it);
    }
  }
  hipLaunchKernelGGL(k_green_res, dim3(rb), dim3(256), 0, st, a, iters);
  return hipGetLastError();
}

// the operator's entries (row, column, value) into the column-major dense
// matrix of the active nodes
__global__ void k_green_scatter(const int *__restrict__ rows, const int *__restrict__ cols,
                                const double *__restrict__ vals, int n, double *__restrict__ A,
                                long long lda) {
  const int e = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (e < n)
    A[(long long)cols[e] * lda + rows[e]] = vals[e];
}

// G[k][j] (row stride ld, FP32) from the lower triangle of the FP64
// inverse of the active nodes; Dirichlet rows and columns and the padding
// are zero
__global__ void k_green_expand(const double *__restrict__ Ainv, long long lda,
                               const int *__restrict__ idx, int nn, int ld,
                               float *__restrict__ G) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long total = (long long)nn * ld;
  if (e >= total)
    return;
  const int k = (int)(e / ld), j = (int)(e - (long long)k * ld);
  float v = 0.0f;
  if (j < nn) {
    const int ak = idx[k], aj = idx[j];
    if (ak >= 0 && aj >= 0) {
      const int hi = ak > aj ? ak : aj, lo = ak > aj ? aj : ak;
      v = (float)Ainv[(long long)lo * lda + hi];
    }
  }
  G[e] = v;
}

hipError_t dev_launch_green_scatter(const int *rows, const int *cols, const double *vals, int n,
                                    double *A, long long lda, hipStream_t st) {
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_green_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rows,
                     cols, vals, n, A, lda);
  return hipGetLastError();
}

hipError_t dev_launch_green_expand(const double *Ainv, long long lda, const int *idx, int nn,
                                   int ld, float *G, hipStream_t st) {
  const long long total = (long long)nn * ld;
  hipLaunchKernelGGL(k_green_expand, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     Ainv, lda, idx, nn, ld, G);
  return hipGetLastError();
}

#define EMCGPU_DEV_INST(R)                                                                   \
  template hipError_t dev_launch_field<R>(const DevFieldArgs<R> *, size_t, bool, hipStream_t);      \
  template hipError_t dev_launch_efield<R>(const DevFieldArgs<R> *, int, hipStream_t);        \
  template hipError_t dev_launch_step<R>(const DevStepArgs<R> &, const DevContactArgs<R> &, int, \
                                         int, size_t, hipStream_t);                           \
  template hipError_t dev_launch_drift<R>(const DevStepArgs<R> &, int, hipStream_t);          \
  template hipError_t dev_launch_contacts<R>(const DevContactArgs<R> &, long long,           \
                                             hipStream_t);                                    \
  template hipError_t dev_launch_reduce<R>(const DevContactArgs<R> &, int, hipStream_t);      \
  template hipError_t dev_launch_init<R>(const DevContactArgs<R> &, const long long *,        \
                                         long long, hipStream_t);                             \
  template hipError_t dev_launch_compact<R>(const DevParticles<R> &, const DevParticles<R> &, \
                                            int *, long long, DevState *, hipStream_t);       \
  template hipError_t dev_launch_zero_r<R>(R *, long long, hipStream_t);                      \
  template int dev_step_blocks_per_cu<R>(size_t);
EMCGPU_DEV_INST(float)
EMCGPU_DEV_INST(double)

}  // namespace emcgpu
