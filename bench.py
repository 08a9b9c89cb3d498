#!/usr/bin/env python3
"""Benchmark: particle-updates/s of the EMC particle step on one MI355X.

Workload (Si bulk, plain form): 10M electrons in the X valleys of Si
(m_l = 0.916, m_t = 0.19, alpha = 0.5/eV, three subvalley pairs), 300 K,
acoustic + six zero-order Jacoboni intervalley phonons (13 tables of 2000
levels on (0, 2] eV), 50 kV/cm along [100], dt = 1 fs.  A "step" advances
all particles by one dt including every scatter event inside it and the
per-step ensemble sums (drift velocity, mean energy, subvalley occupancy).

value      = FP32 (fused steps) particle-updates/s with the state resident
             in HBM, buffers sized before the timed call; the FP64 plain
             kernel (one step per launch) is measured beside it.
roofline   = the FP32 step kernel with the timed call's launch shape: it is
             bound by vector-ALU issue, so `achieved` is the rate of VALU
             lane-instructions on active lanes (SQ_THREAD_CYCLES_VALU per
             particle-update from the committed rocprof counter summary of
             that shape x updates per launch / the launch time measured here
             with HIP events) against the full-rate VALU lane throughput
             (one wave64 instruction per SIMD every 4 cycles, 1024 SIMDs,
             2.4 GHz); the issue rate, the lane utilisation and the HBM view
             (measured traffic, and the 48 B per update equivalent) are
             reported beside it.
cpu_baseline = the oracle's FP64 restatement of the reference loop with the
             reference engine (one mt19937_64 per particle) on the host
             threads this process may use.
Other legs (each a key of the JSON line): the pathwise FP32-vs-FP64
precision statement, the plain 10k-electron run of the CPU reference path
(C1) with the CPU timed beside it, the 64-instance sweep, hot-phonon
cooling (plain and screened) with the bath energy balance, the 2-D
resistor, and CPU legs of those workloads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_FP32 = 48  # k(3) + E + tau + valley|sub, 4 B each, read + write
BYTES_FP64 = 88  # k(3) + E + tau 8 B each + valley|sub 4 B, read + write
# full-rate vector-ALU lane throughput: 1024 SIMDs (256 CUs x 4), one wave64
# instruction per SIMD every 4 cycles, 2.4 GHz peak engine clock -> lane-
# instructions per second (the rate of a stream of v_fma_f32; 32-bit integer
# add/xor streams issue faster, see calib/valu_ceiling.hip)
N_SIMD = 256 * 4
CLOCK_HZ = 2.4e9
VALU_PEAK_TOPS = N_SIMD * 64 * CLOCK_HZ / 4 / 1e12
DT = 1e-15
SEED = 20261202
LEGS = ["headline", "fp64", "fp32_plain", "precision", "c1", "batched", "hot", "hot_screened",
        "device", "cpu"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--particles", type=int, default=10_000_000)
    ap.add_argument("--spl", type=int, default=0, help="FP32 steps fused per launch (0 = auto)")
    ap.add_argument("--ppl", type=int, default=0,
                    help="particles per GPU lane of the production kernel (0 = library default)")
    ap.add_argument("--fp64-steps", type=int, default=0, help="FP64 timed steps (0 = --steps)")
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma-separated subset of: " + ", ".join(LEGS))
    ap.add_argument("--precision-steps", type=int, default=1000)
    ap.add_argument("--batched-steps", type=int, default=2000)
    ap.add_argument("--hot-steps", type=int, default=2000)
    ap.add_argument("--hot-particles", type=int, default=4_000_000)
    ap.add_argument("--device-steps", type=int, default=20000)
    ap.add_argument("--device-particles", type=int, default=2_000_000)
    ap.add_argument("--device-pair-steps", type=int, default=2000,
                    help="steps of the FP32 / FP64 resistor pair")
    ap.add_argument("--cpu-particles", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (rocprof byte-counter calibration)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the vector-ALU issue-rate streams (viennaemc_amd/calib)")
    a = ap.parse_args()
    a.legs = set(x.strip() for x in a.legs.split(",") if x.strip())
    bad = a.legs - set(LEGS)
    if bad:
        ap.error("unknown legs: %s" % ", ".join(sorted(bad)))
    return a


def copy_calibration(mib):
    """One device-to-device copy kernel of known size (torch elementwise copy),
    to scale rocprof's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))


def host_threads():
    """Threads this process may run on: the CPU affinity, capped by
    OMP_NUM_THREADS when set (the GPU host allots 16 per GPU)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def launch_shape(steps, spl):
    """steps fused per launch by the library for a call of `steps` steps"""
    return min(spl if spl > 0 else 64, steps)


def run_gpu(emcgpu, spec, precision, n, warmup, steps, spl, seed=SEED, ppl=0):
    """Thermal init, warm-up, buffers sized for the timed call, then the
    timed call; a second pass of 4 launches of the same shape timed with HIP
    events on the library stream gives the kernel's launch duration."""
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, spec)
    ctx.resize([n])
    ctx.init_particles(0, seed, spec["T"])
    if warmup:
        ctx.step(DT, warmup, seed=seed, step0=0, steps_per_launch=spl, ppl=ppl)
    ctx.reserve(DT, steps, steps_per_launch=spl, ppl=ppl)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(DT, steps, seed=seed, step0=warmup, steps_per_launch=spl, ppl=ppl)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    av = ctx.averages(0)
    K = launch_shape(steps, spl)
    ctx.timing(reset=True)
    ctx.step(DT, 4 * K, seed=seed, step0=warmup + steps, steps_per_launch=K, timing=True, ppl=ppl)
    ms, launches = ctx.timing()
    ctx.close()
    return dict(wall=wall, av=av, kernel_ms=ms / max(launches, 1), launches=launches, K=K)


PROFILE_MAIN = os.path.join(ROOT, "profiles", "emc_step_si_bulk_10M.md")


def profile_figures(path=PROFILE_MAIN):
    """Figures of a committed rocprof summary: table rows `| figure: name |
    value ... |` (VALU lane-instructions per particle-update, calibrated HBM
    bytes per launch of a profiled shape, ...)."""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        parts = [p.strip() for p in line.split("|")]
        if len(parts) >= 4 and parts[1].startswith("figure:"):
            try:
                out[parts[1][len("figure:"):].strip()] = float(parts[2].split()[0])
            except ValueError:
                pass
    return out


# ---------------------------------------------------------------------------
# legs
# ---------------------------------------------------------------------------
def leg_precision(emcgpu, spec, n, steps):
    """FP32 against FP64 on the same Philox stream (pathwise): 16 equal
    groups of the 10M electrons, `steps` steps, time averages over the second
    half; the relative difference of v_d and <E>, its paired standard error
    and the ensemble's own standard error."""
    from viennaemc_amd.workloads import group_stats, run_groups_gpu
    groups = 16
    per = n // groups
    r = {p: run_groups_gpu(emcgpu, spec, p, groups, per, steps, DT, SEED, steps // 2)
         for p in (emcgpu.FP64, emcgpu.FP32)}
    out = {"groups": groups, "particles": per * groups, "steps": steps,
           "window": "steps %d-%d" % (steps // 2, steps - 1)}
    for key, name in (("v", "vd"), ("E", "E")):
        m64, s64 = group_stats(r[0][key])
        md, sd = group_stats(r[1][key] - r[0][key])
        out["delta_%s_rel" % name] = md / abs(m64)
        out["delta_%s_rel_sigma_paired" % name] = sd / abs(m64)
        out["%s_rel_sigma_ensemble" % name] = s64 / abs(m64)
        out["%s_fp64" % name] = m64
    return out


def leg_c1(emcgpu, cpu_threads):
    """The plain bulk Si run of the CPU reference path (10k e-, 10 kV/cm
    [100], 300 K, 3000 steps of 1 fs): GPU FP64 plain and FP32, each in 20
    equal groups for the statistical error, and the oracle's reference-engine
    loop timed on the same workload (step loop only, all host threads);
    observables compared at 3 sigma."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle_py as O
    from viennaemc_amd.workloads import (C1_AVG_FROM, C1_DT, C1_GROUPS, C1_PARTICLES, C1_STEPS,
                                         field_force, group_stats, run_groups_gpu, si_bulk_c1)
    spec = si_bulk_c1()
    per = C1_PARTICLES // C1_GROUPS
    out = {"workload": "Si bulk, 10k e-, 10 kV/cm [100], 300 K, 3000 x 1 fs, %d groups"
                       % C1_GROUPS}
    gpu = {}
    for p, name, spl, ppl in ((emcgpu.FP64, "fp64_plain", 1, 1), (emcgpu.FP32, "fp32", 0, 0)):
        r = run_groups_gpu(emcgpu, spec, p, C1_GROUPS, per, C1_STEPS, C1_DT, SEED, C1_AVG_FROM,
                           steps_per_launch=spl, ppl=ppl)
        gpu[name] = r
        out["gpu_%s_updates_per_s" % name] = C1_PARTICLES * C1_STEPS / r["wall"]
    # CPU: the reference engine's loop, timed on one 10k ensemble
    O.load()
    force, d = field_force(spec)
    Vt = O.thermal_voltage(spec["T"])
    t = np.zeros(1)
    O.OracleModel(spec).run_mt(C1_PARTICLES, SEED, Vt, C1_STEPS, C1_DT, force, d, obs=False,
                               threads=cpu_threads, timing=t)
    out["cpu_updates_per_s"] = C1_PARTICLES * C1_STEPS / t[0]
    out["cpu_threads"] = cpu_threads
    out["cpu_kind"] = "port"

    # CPU observables in equal groups (untimed), streams far apart
    def one(g):
        obs, _, sub = O.OracleModel(spec).run_mt(per, SEED + (g << 32), Vt, C1_STEPS, C1_DT,
                                                  force, d, threads=1, sub_occ=True)
        return np.concatenate([[obs[C1_AVG_FROM:, 1, 0].mean(), obs[C1_AVG_FROM:, 0, 0].mean()],
                               sub[C1_AVG_FROM:, 0, :3].mean(axis=0)])

    with ThreadPoolExecutor(max_workers=cpu_threads) as ex:
        cpu = np.array(list(ex.map(one, range(C1_GROUPS))))
    mc, sc = group_stats(cpu)
    names = ["vd", "E", "occ0", "occ1", "occ2"]
    out["cpu"] = {k: [float(mc[i]), float(sc[i])] for i, k in enumerate(names)}
    for gname, r in gpu.items():
        g = np.column_stack([r["v"], r["E"], r["sub"]])
        mg, sg = group_stats(g)
        z = (mg - mc) / np.sqrt(sg ** 2 + sc ** 2)
        out["gpu_" + gname] = {k: [float(mg[i]), float(sg[i])] for i, k in enumerate(names)}
        out["z_" + gname] = {k: float(z[i]) for i, k in enumerate(names)}
        out["within_3sigma_" + gname] = bool(np.all(np.abs(z) <= 3.0))
    return out


def leg_batched(emcgpu, warmup, steps, with_fp64=True):
    """The 64-instance Si bias/temperature sweep in one arena (tables built
    per instance in FP64, stored in the kinematics' precision), every call
    launched as one HIP graph with the launch shape chosen by emcgpu_autotune
    on this ensemble; whole-sweep particle-updates/s.  With FP64: the same
    sweep in FP64 on the same streams, and per instance the relative
    difference of v_d and <E> over the second half."""
    from viennaemc_amd.workloads import batched_instances

    def run(precision, tune):
        inst = batched_instances(64, SEED)
        counts = [n for _, n in inst]
        ctx = emcgpu.Context(precision=precision, n_instances=len(inst))
        t0 = time.perf_counter()
        for i, (m, _) in enumerate(inst):
            ctx.setup(i, m)
        ctx.resize(counts)
        for i, (m, _) in enumerate(inst):
            ctx.init_particles(i, SEED, m["T"])
        ctx.synchronize()
        res = {"setup_s": time.perf_counter() - t0}
        nl, emax = inst[0][0]["n_levels"], inst[0][0]["max_energy"]
        mats = [dict(rho=m.get("rho", 2329.0), v_sound=m.get("v_sound", 9040.0)) for m, _ in inst]
        if tune:
            t0 = time.perf_counter()
            for i, (m, _) in enumerate(inst):
                ctx.build_tables(i, nl, emax, m["T"], **mats[i])
            res["tables_host_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        ctx.build_tables_all(nl, emax, [m["T"] for m, _ in inst], mats)
        res["tables_gpu_s"] = time.perf_counter() - t0
        if tune:
            t0 = time.perf_counter()
            res["tune"] = ctx.autotune(DT, seed=SEED, step0=0)
            res["tune_s"] = time.perf_counter() - t0
        if warmup:
            ctx.step(DT, warmup, seed=SEED, step0=0, graph=True)
        ctx.reserve(DT, steps)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.step(DT, steps, seed=SEED, step0=warmup, graph=True)
        ctx.synchronize()
        res["wall"] = time.perf_counter() - t0
        half = steps // 2
        res["vd_series"] = np.array([ctx.averages(i)["avgV"][half:, 0] for i in range(len(inst))])
        res["E_series"] = np.array([ctx.averages(i)["avgE"][half:, 0] for i in range(len(inst))])
        res["vd"] = res["vd_series"].mean(axis=1)
        res["E"] = res["E_series"].mean(axis=1)
        res["total"] = sum(counts)
        res["fields"] = np.array([m["field_kvcm"] for m, _ in inst])
        ctx.close()
        return res

    r32 = run(emcgpu.FP32, True)
    tune = r32["tune"]
    out = {"updates_per_s": r32["total"] * steps / r32["wall"], "instances": 64,
           "particles": r32["total"], "steps": steps, "ms_per_step": r32["wall"] / steps * 1e3,
           "setup_s": r32["setup_s"], "tables_host_s": r32["tables_host_s"],
           "tables_gpu_s": r32["tables_gpu_s"],
           "vd_min": float(r32["vd"].min()), "vd_max": float(r32["vd"].max()),
           "autotune": {"particles_per_lane": tune["best"][0], "steps_per_launch": tune["best"][1],
                        "tune_s": r32["tune_s"],
                        "trials_ms_per_step": [[p, k, round(ms, 5)] for p, k, ms in tune["trials"]]},
           "workload": "64 Si bulk instances, field log-U[0.1,200] kV/cm along [100]/[111], "
                       "T U[77,500] K, N log-U[1e4,2e6], FP32, HIP-graph launch, autotuned shape"}
    if with_fp64:
        r64 = run(emcgpu.FP64, False)
        out["fp64_updates_per_s"] = r64["total"] * steps / r64["wall"]
        out["fp32_vs_fp64"] = paired_difference(r32, r64, steps)
    return out


def paired_difference(r32, r64, steps, blocks=10):
    """Per instance: the time-averaged v_d and <E> of FP32 minus FP64 on the
    same Philox streams, relative to FP64, with the standard error of that
    difference from batch means of the per-step difference series (blocks of
    the averaging window; the two runs decorrelate after a draw lands within
    rounding of a decision boundary, so the difference carries sampling
    noise that falls with the instance's particle count)."""
    out = {"window": "steps %d-%d, same Philox streams" % (steps // 2, steps - 1),
           "sigma": "batch means of the per-step difference, %d blocks" % blocks}
    for key, name in (("vd_series", "vd"), ("E_series", "E")):
        d = r32[key] - r64[key]
        ref = np.abs(r64[key].mean(axis=1))
        nb = d.shape[1] // blocks
        bm = d[:, :nb * blocks].reshape(d.shape[0], blocks, nb).mean(axis=2)
        sig = bm.std(axis=1, ddof=1) / np.sqrt(blocks) / ref
        rel = d.mean(axis=1) / ref
        z = rel / np.maximum(sig, 1e-300)
        resolved = sig < 1e-3  # instances whose paired error resolves 0.1 %
        out["delta_%s_rel_max" % name] = float(np.abs(rel).max())
        out["delta_%s_rel_median" % name] = float(np.median(np.abs(rel)))
        out["delta_%s_within_3sigma" % name] = int((np.abs(z) <= 3).sum())
        out["delta_%s_rel_max_resolved" % name] = (float(np.abs(rel[resolved]).max())
                                                  if resolved.any() else None)
        out["instances_resolved_%s" % name] = int(resolved.sum())
    out["instances"] = int(r32["vd_series"].shape[0])
    return out


def leg_hot(emcgpu, n, warmup, steps, precision, screened=False):
    """Photoexcited cooling with a hot-phonon bath: every step is the
    particle step with LDS bath tallies, the bath update and the table
    rebuild (one HIP graph per call); whole-job particle-updates/s, and the
    FP64 energy balance of the timed steps: carrier energy lost = net emitted
    LO energy = LO energy gained + LO energy passed to the acoustic reservoir.
    screened: the hot-carrier default of the reference (screened hot-phonon
    Froehlich, q-resolved rate and angle, plasma screening refreshed after
    every step from the step's tile partials)."""
    from viennaemc_amd.workloads import photoexcited_cooling, screened_cooling
    spec = screened_cooling(n) if screened else photoexcited_cooling(n)
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, spec)
    ctx.resize([n])
    ctx.init_particles_energy(0, SEED, 0.5, 0.05)
    if screened:
        ctx.update_screening(0)
    dt = 0.5e-15
    e0 = float(ctx.get_particles(0)["E"].sum())
    if warmup:
        ctx.step(dt, warmup, seed=SEED, step0=0, graph=True)
        e0 = float(ctx.averages(0)["sumE"][-1, 0])
    ctx.reserve(dt, steps)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(dt, steps, seed=SEED, step0=warmup, graph=True)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    av = ctx.averages(0)
    hist = ctx.bath_history(0, 0)
    en = ctx.bath_energy(0, 0)
    bath = ctx.get_bath(0, 0)
    qs2 = ctx.get_screening(0)
    ctx.timing(reset=True)
    ctx.step(dt, 64, seed=SEED, step0=warmup + steps, timing=True)
    ms, launches = ctx.timing()
    ctx.close()
    out = {"updates_per_s": n * steps / wall, "particles": n, "steps": steps,
           "ms_per_step": wall / steps * 1e3, "step_kernel_ms": ms / max(launches, 1),
           "meanE_first": float(av["avgE"][0, 0]), "meanE_last": float(av["avgE"][-1, 0]),
           "meanNq_last": float(hist[-1])}
    # the energy balance of the timed steps (eV in the bath volume)
    loss = e0 - float(av["sumE"][-1, 0])
    em, lo, dec = (float(x) for x in en.sum(axis=0))
    out["energy_balance_eV"] = {
        "carrier_loss": loss, "net_emitted": em, "lo_gained": lo, "passed_to_reservoir": dec,
        "rel_resid_carrier_vs_emitted": (loss - em) / em,
        "rel_resid_emitted_vs_lo_plus_reservoir": (em - lo - dec) / em}
    if screened:
        out["qs2_last"] = float(qs2)
        out["workload"] = ("screened photoexcited cooling, %d e- N(0.5, 0.05) eV, screened hot-phonon "
                           "Froehlich with q-resolved rate and angle on the 256-bin bath, plasma "
                           "screening (eps 3.5) and window sums refreshed every step, dt 0.5 fs, %s"
                           % (n, "FP32" if precision else "FP64"))
    else:
        out["maxNq_last"] = float(bath["nq"].max())
        out["workload"] = ("photoexcited cooling, %d e- N(0.5, 0.05) eV, Kane Gamma valley, hot-phonon "
                           "Froehlich on a 256-bin LO bath (tau_LO 1 ps into the acoustic reservoir, "
                           "30 ps to the lattice), dt 0.5 fs, bath update + table rebuild every step, "
                           "%s" % (n, "FP32" if precision else "FP64"))
    out["meanE_curve"] = av["avgE"][:, 0]
    return out


def leg_device(emcgpu, n, warmup, steps, precision, seed=SEED):
    """Self-consistent 2-D resistor: 200 x 50 nodes at 1 nm, N_D 1e17 cm^-3
    with 1e19 cm^-3 ohmic regions, 0.5 V, parabolic m* = 0.26 valley with
    acoustic + intervalley phonons, CIC, Poisson (Newton corrections from the
    separable inverse of the discrete operator, multigrid as check and
    fallback, FP64 residual to 1e-10) every 0.1 fs step; whole-job particle-updates/s
    (particles moved per step x steps / wall time, HIP-graph launch)."""
    from viennaemc_amd.workloads import MODELS_DEVICE, resistor_device
    dev = resistor_device(n_particles=n)
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, MODELS_DEVICE["device_para"])
    ctx.device_setup(dev, pm=emcgpu.PM_CIC, poisson_tol=1e-10, poisson_max_iter=50)
    t0 = time.perf_counter()
    n0 = ctx.device_init(seed)
    t_init = time.perf_counter() - t0
    dt = 0.1e-15
    if warmup:
        ctx.device_step(dt, warmup, seed=seed, graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    out = ctx.device_step(dt, steps, seed=seed, graph=True)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    pot = ctx.device_fields()["pot"]
    ctx.close()
    npart = out["n_particles"].astype(np.float64)
    before = np.concatenate([[npart[0] - (out["injected"][0] - out["exits"][0]
                                          - out["deleted"][0]).sum()], npart[:-1]])
    # electrons leaving through each contact per step, net of the contact
    # equilibrium (handleOhmicContacts returns injected - deleted)
    net = (out["exits"] - (out["injected"] - out["deleted"])).astype(np.float64)
    tail = slice(steps // 2, None)
    q = 1.60219e-19
    return {"updates_per_s": before.sum() / wall, "particles_init": int(n0), "steps": steps,
            "ms_per_step": wall / steps * 1e3, "init_s": t_init,
            "particles_last": int(npart[-1]),
            "poisson_newton_iters_mean": float(out["poisson_iters"].mean()),
            "poisson_resid_max": float(out["poisson_resid"].max()),
            "charge_checksum_max": float(out["checksum"].max()),
            "contact_current_A": [float(x) for x in q * net[tail].mean(axis=0) / dt],
            "exits_per_step": float(out["exits"].sum(axis=1).mean()),
            "pot": pot,
            "workload": "2-D n+ n n+ resistor, 200 x 50 nodes at 1 nm, N_D 1e17 / 1e19 cm^-3, "
                        "0.5 V, %d e- (parabolic 0.26 m0, acoustic + IV), CIC, Poisson every "
                        "0.1 fs step, %s" % (n, "FP32" if precision else "FP64")}


def cpu_baseline(spec, n, steps, threads):
    from oracle import oracle_py as O
    from viennaemc_amd.workloads import field_force
    M = O.OracleModel(spec)
    force, d = field_force(spec)
    t = np.zeros(1)
    M.run_mt(n, SEED, O.thermal_voltage(spec["T"]), steps, DT, force, d, obs=False,
             threads=threads, timing=t)
    res = dict(value=n * steps / t[0], unit="particle-updates/s", cores=threads, kind="port",
               sample="%d particles x %d steps of the 10M-electron workload, FP64, mt19937_64 per "
                      "particle, step loop only (%.1f s wall on %d threads)"
                      % (n, steps, t[0], threads))
    cal = profile_figures(os.path.join(ROOT, "profiles", "cpu_calibration_c1.md"))
    if cal:
        # the reference's own handler loop against this port, one thread each on
        # the plain bulk Si run (timed where the reference sources are)
        res["calibration"] = cal
        if cal.get("reference_over_port"):
            res["reference_equivalent_value"] = res["value"] * cal["reference_over_port"]
    return res


def cpu_legs(threads):
    """The CPU side of the other workloads, bounded samples on the host:
    C3 the 64 instances run one after the other (1/100 of each instance's
    particles, 100 steps), C4 the oracle's reference device loop with SOR
    (the resistor with 200k electrons, one thread: the reference's deposit,
    particle removal and Poisson solve are serial), C5 the hot-phonon loop with the bath update
    and table rebuild after every step (400k electrons, 20 steps)."""
    from oracle import oracle_py as O
    from viennaemc_amd.workloads import (MODELS_DEVICE, batched_instances, field_force,
                                         photoexcited_cooling, resistor_device)
    out = {}
    # C3
    tot, t_all = 0, 0.0
    for m, cnt in batched_instances(64, SEED):
        n = max(cnt // 100, 256)
        force, d = field_force(m)
        t = np.zeros(1)
        O.OracleModel(m).run_mt(n, SEED, O.thermal_voltage(m["T"]), 100, DT, force, d, obs=False,
                                threads=threads, timing=t)
        tot += n * 100
        t_all += t[0]
    out["batched_sweep"] = {"updates_per_s": tot / t_all, "threads": threads,
                            "sample": "64 sequential runs, 1/100 of each instance's particles "
                                      "(>= 256) x 100 steps, step loops only"}
    # C4: a 6-step run minus a 1-step run (initialisation and the
    # equilibrium solve cancel)
    dev = resistor_device(n_particles=200_000)
    od = O.OracleDevice(dev)
    t0 = time.perf_counter()
    r1 = od.run(O.OracleModel(MODELS_DEVICE["device_para"]), O.PM_CIC, 1e-8, 1.8, 0.1e-15, 1, SEED)
    t1 = time.perf_counter()
    r6 = od.run(O.OracleModel(MODELS_DEVICE["device_para"]), O.PM_CIC, 1e-8, 1.8, 0.1e-15, 6, SEED)
    t2 = time.perf_counter()
    moved = float(r6["n_part"][:5].sum())
    out["device_resistor"] = {"updates_per_s": moved / ((t2 - t1) - (t1 - t0)), "threads": 1,
                              "particles": int(r1["n_init"][0]),
                              "sample": "oracle device loop with SOR (1e-8 V, omega 1.8) on the "
                                        "200 x 50 resistor with 200k electrons, steps 2-6 (a "
                                        "6-step run minus a 1-step run)"}
    # C5
    n = 400_000
    spec = photoexcited_cooling(n)
    force, d = field_force(spec)
    t = np.zeros(1)
    O.OracleModel(spec).run_mt(n, SEED, O.thermal_voltage(spec["T"]), 20, 0.5e-15, force, d,
                               obs=False, threads=threads, timing=t)
    out["hot_phonon"] = {"updates_per_s": n * 20 / t[0], "threads": threads,
                         "sample": "400k electrons x 20 steps of 0.5 fs with the bath update and "
                                   "table rebuild after every step"}
    return out


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.valu_calibration:
        import subprocess
        return subprocess.call([os.path.join(ROOT, "viennaemc_amd", "calib", "valu_ceiling")])
    if args.gpus != 1:
        print("bench.py measures one GPU: the particle step has no multi-GPU path in this library "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2
    from viennaemc_amd import emcgpu
    from viennaemc_amd.workloads import MODELS
    spec = MODELS["si_jacoboni"]
    n = args.particles
    legs = args.legs
    threads = host_threads()
    out = {}
    res = {
        "metric": "particle-updates/s (Si bulk, 10M e-) FP32-batched vs FP64 plain; % HBM peak; dv_d",
        "value": None, "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (thermal Maxwellian init, Philox-4x32-10)",
    }
    if "headline" in legs:
        r32 = run_gpu(emcgpu, spec, emcgpu.FP32, n, args.warmup, args.steps, args.spl,
                      ppl=args.ppl)
        K = r32["K"]
        t_launch = r32["kernel_ms"] * 1e-3
        res["value"] = n * args.steps / r32["wall"]
        res["ms_per_step"] = r32["wall"] / args.steps * 1e3
        res["config"] = {"workload": "Si bulk X-valley EMC, 10M e-, 50 kV/cm [100], dt 1 fs, 300 K, "
                                     "acoustic + 6 zero-order Jacoboni phonons, 2000-level tables",
                         "particles": n, "steps_per_launch": K,
                         "particles_per_lane": args.ppl or 4,
                         "kernel_ms_per_launch": r32["kernel_ms"]}
        fig = profile_figures()
        upd = n * K  # particle-updates per launch
        hbm_equiv = BYTES_FP32 * upd / t_launch / 1e9
        roof = {"bound": "valu", "achieved": None, "peak": VALU_PEAK_TOPS,
                "unit": "T lane-instr/s (VALU, active lanes)", "frac": None, "traffic": None,
                "hbm_equiv": {"achieved": hbm_equiv, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": hbm_equiv / HBM_PEAK_GBS,
                              "note": "48 B per particle-update x updates per launch / launch time"}}
        lpu = fig.get("valu_lane_instr_per_update_K%d" % K)
        wpu = fig.get("valu_wave_instr_per_update_K%d" % K)
        if lpu:
            roof["achieved"] = lpu * upd / t_launch / 1e12
            roof["frac"] = roof["achieved"] / VALU_PEAK_TOPS
            roof["valu_lane_instr_per_update"] = lpu
        if wpu:
            # wave64 instructions issued per SIMD and cycle at the peak clock
            # (0.25 = one full-rate instruction every 4 cycles)
            roof["valu_wave_instr_per_update"] = wpu
            roof["issue_per_simd_cycle"] = wpu * upd / t_launch / N_SIMD / CLOCK_HZ
            if lpu:
                roof["lane_utilisation"] = lpu / (64 * wpu)
        tb = fig.get("hbm_bytes_per_launch_K%d" % K)
        if tb:
            roof["traffic"] = tb
            roof["hbm_measured"] = {"achieved": tb / t_launch / 1e9, "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": tb / t_launch / 1e9 / HBM_PEAK_GBS}
        res["roofline"] = roof
    if "fp64" in legs:
        s64 = args.fp64_steps or args.steps
        r64 = run_gpu(emcgpu, spec, emcgpu.FP64, n, args.warmup, s64, 1, ppl=1)
        ups64 = n * s64 / r64["wall"]
        out.update(fp64_plain_updates_per_s=ups64, fp64_kernel_ms=r64["kernel_ms"],
                   fp64_plain_hbm_frac=(BYTES_FP64 * n / (r64["kernel_ms"] * 1e-3)) / 1e9
                   / HBM_PEAK_GBS)
        tb = profile_figures().get("hbm_bytes_per_launch_fp64_plain")
        if tb:
            out["fp64_plain_hbm_frac_measured"] = tb / (r64["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
        if res["value"]:
            out["fp32_over_fp64_plain"] = res["value"] / ups64
        r64b = run_gpu(emcgpu, spec, emcgpu.FP64, n, args.warmup, s64, args.spl, ppl=args.ppl)
        out["fp64_batched_updates_per_s"] = n * s64 / r64b["wall"]
    if "fp32_plain" in legs:
        s1 = min(args.steps, 200)
        r1 = run_gpu(emcgpu, spec, emcgpu.FP32, n, 0, s1, 1, ppl=args.ppl)
        out["fp32_plain_updates_per_s"] = n * s1 / r1["wall"]
        out["fp32_plain_kernel_ms"] = r1["kernel_ms"]
        out["fp32_plain_hbm_frac"] = BYTES_FP32 * n / (r1["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
        tb = profile_figures().get("hbm_bytes_per_launch_K1")
        if tb:
            # measured traffic (state + partial sums) of the same launch shape
            out["fp32_plain_hbm_frac_measured"] = (tb / (r1["kernel_ms"] * 1e-3) / 1e9
                                                   / HBM_PEAK_GBS)
    if "precision" in legs:
        out["precision_fp32_vs_fp64"] = leg_precision(emcgpu, spec, n, args.precision_steps)
    if "c1" in legs:
        out["c1"] = leg_c1(emcgpu, threads)
    if "batched" in legs:
        out["batched_sweep"] = leg_batched(emcgpu, args.warmup, args.batched_steps)
    if "hot" in legs:
        h32 = leg_hot(emcgpu, args.hot_particles, args.warmup, args.hot_steps, emcgpu.FP32)
        h64 = leg_hot(emcgpu, args.hot_particles, args.warmup, args.hot_steps, emcgpu.FP64)
        c32, c64 = h32.pop("meanE_curve"), h64.pop("meanE_curve")
        h32["fp32_vs_fp64_meanE_rel_max"] = float(np.abs(c32 - c64).max() / np.abs(c64).max())
        out["hot_phonon"], out["hot_phonon_fp64"] = h32, h64
    if "hot_screened" in legs:
        s32 = leg_hot(emcgpu, args.hot_particles, args.warmup, args.hot_steps, emcgpu.FP32,
                      screened=True)
        s64 = leg_hot(emcgpu, args.hot_particles, args.warmup, args.hot_steps, emcgpu.FP64,
                      screened=True)
        c32, c64 = s32.pop("meanE_curve"), s64.pop("meanE_curve")
        s32["fp32_vs_fp64_meanE_rel_max"] = float(np.abs(c32 - c64).max() / np.abs(c64).max())
        s32["fp32_vs_fp64_qs2_rel"] = abs(s32["qs2_last"] - s64["qs2_last"]) / s64["qs2_last"]
        out["hot_phonon_screened"], out["hot_phonon_screened_fp64"] = s32, s64
    if "device" in legs:
        d = leg_device(emcgpu, args.device_particles, args.warmup, args.device_steps, emcgpu.FP32)
        d.pop("pot")
        out["device_resistor"] = d
        # FP32 and FP64 on equal step counts from the same seed
        p32 = leg_device(emcgpu, args.device_particles, args.warmup, args.device_pair_steps,
                         emcgpu.FP32)
        p64 = leg_device(emcgpu, args.device_particles, args.warmup, args.device_pair_steps,
                         emcgpu.FP64)
        pot32, pot64 = p32.pop("pot"), p64.pop("pot")
        out["device_resistor_fp64"] = p64
        out["device_resistor_fp32_vs_fp64"] = {
            "steps": args.device_pair_steps,
            "fp32_updates_per_s": p32["updates_per_s"],
            "contact_current_A_fp32": p32["contact_current_A"],
            "contact_current_A_fp64": p64["contact_current_A"],
            "potential_max_abs_diff_rel": float(np.abs(pot32 - pot64).max() / np.abs(pot64).max()),
            "particles_last_fp32": p32["particles_last"], "particles_last_fp64": p64["particles_last"]}
    if "cpu" in legs:
        res["cpu_baseline"] = cpu_baseline(spec, args.cpu_particles, args.cpu_steps, threads)
        out["cpu_legs"] = cpu_legs(threads)
    res.update(out)
    print(json.dumps(res, default=float), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
